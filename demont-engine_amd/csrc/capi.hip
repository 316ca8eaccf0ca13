// C ABI of the gfx950 path tracer (declared in include/dmt.h).
//
// The device owns one HIP stream; a scene owns its HBM buffers (acceleration
// data, wavefront queues, accumulation buffer).  Every call validates its
// arguments and reports failures through a status code and dmt_last_error(),
// as the reference's software device reports through a false/0 return and
// LOG_ERROR (src/rhi_software/SoftwareDevice.cpp:337-421, 539-558).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/dmt.h"
#include "host/prep.h"
#include "kernels/adjoint.h"
#include "kernels/output.h"
#include "kernels/lbvh.h"
#include "kernels/refit.h"
#include "kernels/persistent.h"
#include "kernels/wavefront.h"

using namespace dmt;

// Frames of per-frame dmt_render calls kept in flight (one stream each).
// Default frame schedule: the whole frame, camera rays included, by one
// k_tail launch (measured faster than the camera bounce by the per-bounce
// kernels on both the 1080p mesh and the many-light scene).
constexpr uint32_t kDefaultTailAfter = 0;
constexpr uint32_t kTailBlocksPerCu = kTailWavesPerSimd;   // k_tail grid: resident workgroups per CU
constexpr uint32_t kExtendBlocksPerCu = 8;   // k_extend grid: 8 waves per SIMD, rays handed out dynamically
#ifndef DMT_FRAME_SETS
#define DMT_FRAME_SETS 4
#endif
// Frames in flight, one HIP stream each (the runtime gives a process four
// hardware queues by default).
constexpr uint32_t kFrameSets = DMT_FRAME_SETS;

namespace {

thread_local std::string g_err;

dmt_status fail(dmt_status s, const std::string& msg) {
    g_err = msg;
    return s;
}

dmt_status hip_fail(hipError_t e, const char* what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? DMT_ERR_OUT_OF_MEMORY : DMT_ERR_DEVICE;
}

template <class T>
hipError_t dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc((void**)p, count * sizeof(T));
}

}  // namespace

// Scratch of dmt_trace_closest_noalloc: the traversal-stack spill of its
// fixed grid (DMT_NOALLOC_BLOCKS workgroups of kBlock threads, each thread
// DMT_NOALLOC_SPILL_DEPTH entries past its LDS stack), 24 MiB in every GPU's
// copy of this code object, held until the process ends.  In host code the
// name is a stand-in, not a device address: the device's own address comes
// from hipGetSymbolAddress (dmt_device::noalloc_scratch).
static_assert(DMT_NOALLOC_BLOCKS * 256u * DMT_NOALLOC_SPILL_DEPTH * 4u == 24u << 20, "scratch size");
static __device__ __align__(256) uint32_t scratch[(size_t)DMT_NOALLOC_BLOCKS * kBlock * DMT_NOALLOC_SPILL_DEPTH];

struct dmt_device {
    int id = 0;
    uint32_t* noalloc_scratch = nullptr;   // this GPU's `scratch` (null: the lookup failed, noalloc_error says why)
    std::string noalloc_error;
    hipStream_t stream = nullptr;   // exported stream; frame set 0, uploads, queries
    hipStream_t aux[kFrameSets - 1] = {};   // frame sets 1.. (frames in flight)
    bool lost = false;
    // Scenes hold the device: dmt_device_destroy with live scenes defers the
    // teardown to the last dmt_scene_destroy.
    uint32_t scenes = 0;
    bool destroy_requested = false;
    uint32_t cus = 0;               // compute units
};

namespace {
void device_free(dmt_device* d) {
    (void)hipSetDevice(d->id);
    hipStream_t all[kFrameSets] = {d->stream};
    for (uint32_t k = 1; k < kFrameSets; ++k) all[k] = d->aux[k - 1];
    for (hipStream_t st : all) {
        if (!st) continue;
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    delete d;
}
}  // namespace

struct dmt_scene {
    dmt_device* dev = nullptr;
    SceneView view{};
    std::vector<void*> buffers;       // scene-lifetime device allocations
    uint32_t bvh_depth = 0;
    uint32_t bvh_nodes = 0;
    uint64_t scene_bytes = 0;
    // kFrameSets wavefront sets: frame k of per-frame calls runs on set
    // k % kFrameSets and its stream, so the latency-bound end of one frame
    // (the last paths of its k_tail launch) overlaps the next frames' camera
    // bounce.  Only the final accumulation is ordered frame after frame.
    Wavefront wf[kFrameSets]{};
    std::vector<void*> wf_buffers;
    float4* accum = nullptr;
    uint32_t acc_w = 0, acc_h = 0;
    bool accum_external = false;      // caller-owned (dmt_scene_bind_accum)
    DeviceStats* stats = nullptr;
    DeviceStats* count_stats = nullptr;
    hipEvent_t acc_done[kFrameSets] = {};
    hipEvent_t poll[kFrameSets][kPollSlots] = {};
    uint32_t* host_count = nullptr;   // pinned, kFrameSets * kPollSlots
    int next_set = 0;
    int last_set = -1;
    // wavefront iterations of a frame before k_tail completes it
    uint32_t tail_after = kDefaultTailAfter;
    // profiling
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<uint8_t> ev_kind;   // per event pair: 0 k_extend, 1 k_tail
    double extend_ms = 0.0;
    uint64_t extend_launches = 0;
    double tail_ms = 0.0;
    uint64_t tail_launches = 0;
    uint64_t frames = 0;
    // Batches hold their scenes: dmt_scene_destroy of a member is deferred
    // to the last dmt_batch_destroy, and the resolution is fixed meanwhile.
    uint32_t batches = 0;
    bool destroy_requested = false;
    // material parameters: host copy of the prims and the device array
    std::vector<float> prims_host;
    float4* prims_dev = nullptr;
    // mesh refit (dmt_scene_update_mesh): vertex indices per leaf slot, the
    // triangle BVH's internal nodes by level (deepest level first), the
    // device copy of the positions
    uint32_t mesh_vertices = 0;
    uint32_t mesh_tris = 0;
    uint32_t* leaf_idx = nullptr;
    uint32_t* refit_nodes = nullptr;
    std::vector<uint32_t> refit_level_start;   // offsets into refit_nodes, one past the end last
    float* positions_dev = nullptr;
    uint32_t* adj_spill = nullptr;    // traversal spill area of the adjoint kernel
    float4* adj_rec = nullptr;        // the adjoint's metal-vertex records and their count
    size_t adj_rec_bytes = 0;
    uint32_t* adj_rec_count = nullptr;
    float4* moon = nullptr;           // moon texture (dmt_scene_set_moon_map)
    uchar4* atlas = nullptr;          // PBR atlas (dmt_scene_set_pbr_atlas)
    bool prims_textured = false;      // an analytic primitive has a tile the atlas holds
    bool has_water = false;           // a primitive or cluster of material 3
};

namespace {

dmt_status check_device(dmt_device* dev) {
    if (!dev) return fail(DMT_ERR_INVALID_ARG, "device is null");
    if (dev->lost) return fail(DMT_ERR_DEVICE_LOST, "device lost");
    hipError_t e = hipSetDevice(dev->id);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    return DMT_OK;
}

dmt_status mark_lost_if_fatal(dmt_device* dev, hipError_t e, const char* what) {
    if (e == hipErrorLaunchFailure || e == hipErrorIllegalAddress || e == hipErrorAssert) dev->lost = true;
    return hip_fail(e, what);
}

// Copies count elements to a new device buffer followed by `pad` zero
// elements (leaf primitive arrays: a traversal fetch may read past the last
// primitive of a leaf).
template <class T>
dmt_status upload(dmt_scene* s, const T* src, size_t count, const T** out, size_t pad = 0) {
    *out = nullptr;
    if (count == 0) return DMT_OK;
    T* d = nullptr;
    hipError_t e = dalloc(&d, count + pad);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(scene)");
    s->buffers.push_back(d);
    s->scene_bytes += (count + pad) * sizeof(T);
    e = hipMemcpy(d, src, count * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(scene)");
    if (pad) {
        e = hipMemset(d + count, 0, pad * sizeof(T));
        if (e != hipSuccess) return hip_fail(e, "hipMemset(scene)");
    }
    *out = d;
    return DMT_OK;
}

hipStream_t set_stream(const dmt_scene* s, int set) { return set == 0 ? s->dev->stream : s->dev->aux[set - 1]; }

hipError_t sync_all(dmt_device* d) {
    hipError_t e = hipStreamSynchronize(d->stream);
    for (uint32_t k = 0; k + 1 < kFrameSets; ++k) {
        hipError_t e2 = hipStreamSynchronize(d->aux[k]);
        if (e == hipSuccess) e = e2;
    }
    return e;
}

void free_sets(dmt_scene* s) {
    for (void* p : s->wf_buffers) (void)hipFree(p);
    s->wf_buffers.clear();
    for (auto& w : s->wf) w = Wavefront{};
    s->last_set = -1;
    s->next_set = 0;
}

void free_accum(dmt_scene* s) {
    free_sets(s);
    if (s->accum && !s->accum_external) (void)hipFree(s->accum);
    s->accum = nullptr;
    s->accum_external = false;
    s->acc_w = s->acc_h = 0;
}

// The accumulation buffer at w x h (zeroed when it is (re)allocated).
dmt_status ensure_accum(dmt_scene* s, uint32_t w, uint32_t h, bool* resized) {
    *resized = false;
    if (s->accum && s->acc_w == w && s->acc_h == h) return DMT_OK;
    if (s->batches > 0) return fail(DMT_ERR_INVALID_ARG, "the scene belongs to a batch at another resolution");
    if (s->accum_external)
        return fail(DMT_ERR_INVALID_ARG, "the bound accumulation buffer is " + std::to_string(s->acc_w) + "x" +
                                             std::to_string(s->acc_h));
    hipError_t e = sync_all(s->dev);
    if (e != hipSuccess) return mark_lost_if_fatal(s->dev, e, "synchronize before resize");
    free_accum(s);
    *resized = true;
    const size_t n = (size_t)w * h;
    e = hipMalloc((void**)&s->accum, n * 16);
    if (e != hipSuccess) {
        s->accum = nullptr;
        return hip_fail(e, "hipMalloc(accum)");
    }
    e = hipMemsetAsync(s->accum, 0, n * 16, s->dev->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->dev->stream);
    if (e != hipSuccess) {
        free_accum(s);
        return hip_fail(e, "hipMemsetAsync(accum)");
    }
    s->acc_w = w;
    s->acc_h = h;
    return DMT_OK;
}

// The accumulation buffer and the wavefront sets at w x h.
// The shadow queues of the sky-light records (sun or HDRI cluster, moon) in
// every frame set, allocated when a frame first has sky lights.
dmt_status ensure_sky_queues(dmt_scene* s) {
    if (s->wf[0].sq[2].o) return DMT_OK;
    const size_t n = s->wf[0].capacity;
    for (int set = 0; set < (int)kFrameSets; ++set) {
        for (int k = 2; k < 4; ++k) {
            ShadowQueue& q = s->wf[set].sq[k];
            void** parts[] = {(void**)&q.o, (void**)&q.d, (void**)&q.a, (void**)&q.b, (void**)&q.w};
            for (int i = 0; i < 5; ++i) {
                hipError_t e = hipMalloc(parts[i], i < 4 ? n * 16 : n * 4);
                if (e != hipSuccess) {
                    *parts[i] = nullptr;
                    free_sets(s);
                    return hip_fail(e, "hipMalloc(sky-light shadow queues)");
                }
                s->wf_buffers.push_back(*parts[i]);
            }
        }
    }
    return DMT_OK;
}

dmt_status ensure_wavefront(dmt_scene* s, uint32_t w, uint32_t h, bool* resized) {
    dmt_status st = ensure_accum(s, w, h, resized);
    if (st) return st;
    const size_t n = (size_t)w * h;
    if (s->wf[0].capacity == n) return DMT_OK;
    free_sets(s);
    std::vector<std::pair<void**, size_t>> req;
    for (int set = 0; set < (int)kFrameSets; ++set) {
        Wavefront& W = s->wf[set];
        W.capacity = (uint32_t)n;
        const uint32_t need = 3u * s->bvh_depth + 4u;
        W.spill_depth = need > kLdsDepth ? need - kLdsDepth : 0u;
        W.spill_threads = kMaxGridBlocks * kBlock;
        W.stats = s->stats;
        W.host_count = s->host_count + set * kPollSlots;
        const std::pair<void**, size_t> per_set[] = {
            {(void**)&W.ps.T, n * 16}, {(void**)&W.ps.L, n * 16}, {(void**)&W.ps.F, n * 16},
            {(void**)&W.ps.rng, n * 4}, {(void**)&W.rq[0].o, n * 16}, {(void**)&W.rq[0].d, n * 16},
            {(void**)&W.rq[1].o, n * 16}, {(void**)&W.rq[1].d, n * 16}, {(void**)&W.hits, n * 8},
            {(void**)&W.counters, kCounterSlotsAll * 4}, {(void**)&W.ps.meta, n * 4},
            {(void**)&W.finished, n * 4}, {(void**)&W.regen, n}, {(void**)&W.ps.cost, n * 4},
            {(void**)&W.tile_bucket, (size_t)tile_capacity((uint32_t)n) * 4},
            {(void**)&W.tile_order, (size_t)tile_capacity((uint32_t)n) * 4},
        };
        for (const auto& r : per_set) req.push_back(r);
        for (int k = 0; k < 2; ++k) {
            req.push_back({(void**)&W.sq[k].o, n * 16});
            req.push_back({(void**)&W.sq[k].d, n * 16});
            req.push_back({(void**)&W.sq[k].a, n * 16});
            req.push_back({(void**)&W.sq[k].b, n * 16});
            req.push_back({(void**)&W.sq[k].w, n * 4});
        }
        if (W.spill_depth) req.push_back({(void**)&W.spill, (size_t)W.spill_threads * W.spill_depth * 4});
    }
    for (auto& r : req) {
        hipError_t e = hipMalloc(r.first, r.second);
        if (e != hipSuccess) {
            *r.first = nullptr;
            free_sets(s);
            return hip_fail(e, "hipMalloc(wavefront)");
        }
        s->wf_buffers.push_back(*r.first);
    }
    hipError_t e = hipSuccess;
    for (int set = 0; set < (int)kFrameSets && e == hipSuccess; ++set) {
        e = hipMemsetAsync(s->wf[set].regen, 0, n, s->dev->stream);
        if (e == hipSuccess) e = hipMemsetAsync(s->wf[set].ps.cost, 0, n * 4, s->dev->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->dev->stream);
    if (e != hipSuccess) {
        free_sets(s);
        return hip_fail(e, "hipMemsetAsync");
    }
    return DMT_OK;
}

dmt_status check_frame(const dmt_frame_params* f) {
    if (!f) return fail(DMT_ERR_INVALID_ARG, "frame params are null");
    if (f->width == 0 || f->height == 0) return fail(DMT_ERR_INVALID_ARG, "width and height must be > 0");
    if ((uint64_t)f->width * f->height > (1ull << 26)) return fail(DMT_ERR_LIMIT, "more than 2^26 pixels");
    if (f->width > 65535u || f->height > 65535u) return fail(DMT_ERR_LIMIT, "width and height must be <= 65535");
    if (f->spp < 1 || f->spp > DMT_MAX_SPP) return fail(DMT_ERR_INVALID_ARG, "spp must be in 1..32");
    if (f->max_bounces > DMT_MAX_BOUNCES) return fail(DMT_ERR_LIMIT, "max_bounces must be <= 64");
    if (f->refract_bounces > 64) return fail(DMT_ERR_LIMIT, "refract_bounces must be <= 64");
    if (f->flags & ~(DMT_FRAME_NO_RUSSIAN_ROULETTE | DMT_FRAME_FIXED_JITTER | DMT_FRAME_MESH_TEXTURES))
        return fail(DMT_ERR_INVALID_ARG, "unknown frame flags");
    if ((f->flags & DMT_FRAME_FIXED_JITTER) &&
        !(std::fabs(f->jitter[0]) <= 0.5f && std::fabs(f->jitter[1]) <= 0.5f))
        return fail(DMT_ERR_INVALID_ARG, "fixed jitter must be in [-0.5, 0.5]");
    const float* sky_f[] = {f->sun, f->sun_extra, f->moon, f->moon_extra};
    for (const float* v : sky_f)
        for (int a = 0; a < 4; ++a)
            if (!std::isfinite(v[a])) return fail(DMT_ERR_INVALID_ARG, "sun and moon parameters must be finite");
    if (!std::isfinite(f->sun_disc_brightness)) return fail(DMT_ERR_INVALID_ARG, "sun_disc_brightness must be finite");
    if (f->hdri_light_count > kMaxHdriLights) return fail(DMT_ERR_LIMIT, "at most 8 HDRI lights");
    for (uint32_t i = 0; i < f->hdri_light_count; ++i)
        for (int a = 0; a < 4; ++a)
            if (!std::isfinite(f->hdri_dir[i][a]) || !std::isfinite(f->hdri_col[i][a]))
                return fail(DMT_ERR_INVALID_ARG, "HDRI light values must be finite");
    const float lens[6] = {f->dof[0], f->dof[1], f->dof[2], f->dof[3], f->cat_eye, f->cauchy_c};
    for (float v : lens)
        if (!std::isfinite(v)) return fail(DMT_ERR_INVALID_ARG, "lens and spectral parameters must be finite");
    if (f->dof[0] > 0.0f && f->dof[2] > 64.0f) return fail(DMT_ERR_LIMIT, "at most 64 aperture blades");
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(f->mesh_shift_prev[a]) || !std::isfinite(f->mesh_shift_curr[a]))
            return fail(DMT_ERR_INVALID_ARG, "mesh translations must be finite");
    if (f->motion_blur > 1u || f->spectral > 1u || f->mesh_shift_active > 1u)
        return fail(DMT_ERR_INVALID_ARG, "motion_blur, spectral and mesh_shift_active are 0 or 1");
    if (f->reserved) return fail(DMT_ERR_INVALID_ARG, "reserved must be 0");
    const float water[11] = {f->water_absorb[0], f->water_absorb[1], f->water_absorb[2], f->water_ior, f->wave[0],
                             f->wave[1], f->wave[2], f->wave[3], f->deep_color[0], f->deep_color[1], f->deep_color[2]};
    for (float v : water)
        if (!std::isfinite(v)) return fail(DMT_ERR_INVALID_ARG, "water parameters must be finite");
    for (int a = 0; a < 3; ++a)
        if (f->water_absorb[a] < 0.0f || f->deep_color[a] < 0.0f)
            return fail(DMT_ERR_INVALID_ARG, "water absorption and deep colour must be >= 0");
    if (f->camera_in_water > 1u) return fail(DMT_ERR_INVALID_ARG, "camera_in_water is 0 or 1");
    return DMT_OK;
}

// check_frame, and the frame's fields that the scene makes read: the SDF
// sphere-trace parameters of a scene with clusters.
dmt_status check_frame_for(const dmt_scene* s, const dmt_frame_params* f) {
    dmt_status st = check_frame(f);
    if (st) return st;
    if (s->view.sdf_count) {
        if (f->sdf_max_iters < 1u || f->sdf_max_iters > 256u)
            return fail(DMT_ERR_INVALID_ARG, "sdf_max_iters must be in 1..256 for a scene with SDF clusters");
        if (!(std::isfinite(f->sdf_epsilon) && f->sdf_epsilon > 0.0f))
            return fail(DMT_ERR_INVALID_ARG, "sdf_epsilon must be finite and > 0 for a scene with SDF clusters");
    }
    return DMT_OK;
}

FrameArgs frame_args(const dmt_frame_params* f) {
    FrameArgs F = make_frame_args(f->pos_fovtan, f->fwd_aspect, f->right_xyz, f->up_xyz, f->width, f->height,
                                  f->frame_index, f->spp, f->max_bounces, f->reset_accum, f->mis_enabled,
                                  f->light_tree_enabled, f->refract_bounces, f->firefly_clamp, f->ema_alpha,
                                  f->flags);
    F.cauchy_c = f->cauchy_c;
    return F;
}

// Sky lights that change the frame: with an env map its HDRI cluster lights;
// without one a sun or moon whose disc can show (direction set, y >= -0.1:
// sunDisc / moonDisc are 0 below, and a zero direction meets no ray).
bool sky_lights(const dmt_scene* s, const dmt_frame_params* f) {
    if (s->view.env_present) return f->hdri_light_count > 0u;
    auto up = [](const float* d) { return (d[0] != 0.0f || d[1] != 0.0f || d[2] != 0.0f) && !(d[1] < -0.10f); };
    return up(f->sun) || up(f->moon);
}

// A frame with water: the scene has water surfaces, or the camera starts in
// the water.
bool water_active(const dmt_scene* s, const dmt_frame_params* f) { return s->has_water || f->camera_in_water; }

// The mesh's map tiles a frame asks for (none without DMT_FRAME_MESH_TEXTURES).
void mesh_tex_of(const dmt_frame_params* f, uint32_t out[4]) {
    for (int k = 0; k < 4; ++k) out[k] = (f->flags & DMT_FRAME_MESH_TEXTURES) ? f->mesh_tex[k] : kPbrNoTile;
}

// A frame that samples the PBR atlas: an analytic primitive has a tile the
// atlas holds, or the frame gives the UV-mapped mesh one.
bool pbr_active(const dmt_scene* s, const dmt_frame_params* f) {
    if (!s->view.atlas) return false;
    if (s->prims_textured) return true;
    if (!s->view.mesh_uv) return false;
    uint32_t t[4];
    mesh_tex_of(f, t);
    for (uint32_t k : t)
        if (k < s->view.atlas_tiles) return true;
    return false;
}

// The kernel variant a frame runs: kExtSky with its sky lights (and the four
// shadow-record kinds), kExtSdf with the scene's SDF clusters, kExtPbr with
// its material maps, kExtWater with water.
uint32_t ext_bits(const dmt_scene* s, const dmt_frame_params* f) {
    return (sky_lights(s, f) ? kExtSky : 0u) | (s->view.sdf_count > 0u ? kExtSdf : 0u) |
           (pbr_active(s, f) ? kExtPbr : 0u) | (water_active(s, f) ? kExtWater : 0u);
}

// A frame that draws other numbers than the jitter per sample (lens point,
// shutter time, hero wavelength), takes a fixed jitter, translates the mesh,
// has sky lights, samples the PBR atlas or has water, and any frame of a scene
// with SDF clusters, renders through k_tail<COUNT, true>: its traversal and
// shading are the ones that see them.
bool uses_features(const dmt_scene* s, const dmt_frame_params* f) {
    return f->dof[0] > 0.0f || f->motion_blur || f->spectral || f->mesh_shift_active ||
           (f->flags & DMT_FRAME_FIXED_JITTER) || sky_lights(s, f) || s->view.sdf_count > 0u || pbr_active(s, f) ||
           water_active(s, f);
}

FeatureArgs feature_args(const dmt_scene* s, const dmt_frame_params* f) {
    FeatureArgs X{};
    X.fixed_jitter = (f->flags & DMT_FRAME_FIXED_JITTER) ? 1u : 0u;
    X.jit_x = f->jitter[0];
    X.jit_y = f->jitter[1];
    X.ext = ext_bits(s, f);
    X.cam_in_water = f->camera_in_water ? 1u : 0u;
    if (sky_lights(s, f)) {
        SkyArgs& K = X.sky_args;
        K.sun = f3{f->sun[0], f->sun[1], f->sun[2]};
        K.sun_size = f->sun_extra[0];
        K.sun_dist = f->sun_extra[1];
        K.sun_vflat = f->sun_extra[2];
        K.sun_bright = f->sun_disc_brightness;
        K.moon = f3{f->moon[0], f->moon[1], f->moon[2]};
        K.moon_phase = f->moon[3];
        K.moon_size = f->moon_extra[0];
        K.moon_dist = f->moon_extra[1];
        K.moon_bright = f->moon_extra[2];
        K.hdri_count = s->view.env_present ? f->hdri_light_count : 0u;
        for (uint32_t i = 0; i < kMaxHdriLights; ++i) {
            K.hdri_dir[i] = make_float4(f->hdri_dir[i][0], f->hdri_dir[i][1], f->hdri_dir[i][2], f->hdri_dir[i][3]);
            K.hdri_col[i] = make_float4(f->hdri_col[i][0], f->hdri_col[i][1], f->hdri_col[i][2], f->hdri_col[i][3]);
        }
    }
    X.aperture = f->dof[0];
    X.focal = f->dof[1];
    // int(dof_params.z + 0.5) (PathTrace.slang:5789); below 3 every count is
    // the round aperture, and check_frame bounds it above
    const float blades = f->dof[2] + 0.5f;
    X.blades = (blades >= 3.0f && blades < 65.0f) ? (int)blades : 0;
    X.rot = f->dof[3];
    X.cat_eye = f->cat_eye;
    X.motion_blur = f->motion_blur;
    X.spectral = f->spectral;
    return X;
}

// The scene as a frame sees it: the mesh translation and the SDF sphere-trace
// parameters are per frame.
SceneView frame_view(const dmt_scene* s, const dmt_frame_params* f) {
    SceneView V = s->view;
    V.sdf_iters = f->sdf_max_iters;
    V.sdf_eps = f->sdf_epsilon;
    V.mesh_shift = f->mesh_shift_active;
    V.mesh_prev = f3{f->mesh_shift_prev[0], f->mesh_shift_prev[1], f->mesh_shift_prev[2]};
    V.mesh_curr = f3{f->mesh_shift_curr[0], f->mesh_shift_curr[1], f->mesh_shift_curr[2]};
    mesh_tex_of(f, V.mesh_tex);
    V.water_absorb = f3{f->water_absorb[0], f->water_absorb[1], f->water_absorb[2]};
    V.water_ior = f->water_ior;
    V.wave = make_float4(f->wave[0], f->wave[1], f->wave[2], f->wave[3]);
    V.deep_color = f3{f->deep_color[0], f->deep_color[1], f->deep_color[2]};
    // the ray cone's spread: one pixel's angle, 2 tan(fov_y / 2) / height
    // (PathTrace.slang:5918)
    V.cone_spread = (2.0f * f->pos_fovtan[3]) / (float)std::max(f->height, 1u);
    return V;
}

dmt_status harvest_events(dmt_scene* s) {
    if (s->ev_used == 0) return DMT_OK;
    hipError_t e = sync_all(s->dev);
    if (e != hipSuccess) return mark_lost_if_fatal(s->dev, e, "synchronize");
    for (size_t k = 0; k + 1 < s->ev_used; k += 2) {
        float ms = 0.0f;
        const bool tail = s->ev_kind[k / 2] != 0u;
        if (hipEventElapsedTime(&ms, s->ev_pool[k], s->ev_pool[k + 1]) == hipSuccess) (tail ? s->tail_ms : s->extend_ms) += ms;
        (tail ? s->tail_launches : s->extend_launches)++;
    }
    s->ev_used = 0;
    return DMT_OK;
}

}  // namespace

extern "C" {

const char* dmt_last_error(void) { return g_err.c_str(); }

dmt_status dmt_get_abi_version(uint32_t* v) {
    if (!v) return fail(DMT_ERR_INVALID_ARG, "null pointer");
    *v = DMT_ABI_VERSION;
    return DMT_OK;
}

dmt_status dmt_device_create(int32_t index, dmt_device** out) {
    if (!out) return fail(DMT_ERR_INVALID_ARG, "out is null");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(DMT_ERR_NO_DEVICE, "no HIP device");
    if (index < 0 || index >= count) return fail(DMT_ERR_NO_DEVICE, "device index out of range");
    hipError_t e = hipSetDevice(index);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    auto* d = new (std::nothrow) dmt_device();
    if (!d) return fail(DMT_ERR_OUT_OF_MEMORY, "host allocation");
    d->id = index;
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, index);
    if (e != hipSuccess || cus <= 0) {
        delete d;
        return hip_fail(e, "hipDeviceGetAttribute(CU count)");
    }
    d->cus = (uint32_t)cus;
    e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
    for (uint32_t k = 0; k + 1 < kFrameSets && e == hipSuccess; ++k)
        e = hipStreamCreateWithFlags(&d->aux[k], hipStreamNonBlocking);
    if (e != hipSuccess) {
        device_free(d);
        return hip_fail(e, "hipStreamCreate");
    }
    // this GPU's copy of the allocation-free query's scratch (device `index`
    // is current); only that entry point needs it, so a failed lookup is
    // reported there
    e = hipGetSymbolAddress((void**)&d->noalloc_scratch, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        d->noalloc_scratch = nullptr;
        d->noalloc_error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        (void)hipGetLastError();
    }
    *out = d;
    return DMT_OK;
}

void dmt_device_destroy(dmt_device* d) {
    if (!d) return;
    if (d->scenes > 0) {
        d->destroy_requested = true;
        return;
    }
    device_free(d);
}

dmt_status dmt_device_stream(dmt_device* d, void** out) {
    if (!d || !out) return fail(DMT_ERR_INVALID_ARG, "null pointer");
    *out = (void*)d->stream;
    return DMT_OK;
}

dmt_status dmt_device_synchronize(dmt_device* d) {
    dmt_status st = check_device(d);
    if (st) return st;
    hipError_t e = sync_all(d);
    if (e != hipSuccess) return mark_lost_if_fatal(d, e, "hipStreamSynchronize");
    return DMT_OK;
}

int32_t dmt_device_is_lost(const dmt_device* d) { return d && d->lost ? 1 : 0; }

dmt_status dmt_scene_create(dmt_device* dev, const dmt_scene_desc* desc, const dmt_build_params* bp,
                            dmt_scene** out) {
    if (!out) return fail(DMT_ERR_INVALID_ARG, "out is null");
    *out = nullptr;
    dmt_status st = check_device(dev);
    if (st) return st;
    if (!desc) return fail(DMT_ERR_INVALID_ARG, "scene desc is null");
    if (bp && (bp->flags & ~DMT_BUILD_DEVICE)) return fail(DMT_ERR_INVALID_ARG, "unknown build flags");
    const bool device_build = bp && (bp->flags & DMT_BUILD_DEVICE);
    PreparedScene P;
    std::string msg;
    try {
        EnvCdfInput cdf;
        cdf.marginal = desc->env_marginal;
        cdf.conditional = desc->env_conditional;
        cdf.total = desc->env_total_luminance;
        msg = prepare_scene(desc->prims, desc->prim_count, desc->lights, desc->light_count, desc->positions,
                            desc->vertex_count, desc->indices, desc->index_count, desc->env_rgb, desc->env_width,
                            desc->env_height, desc->env_intensity, bp ? bp->sah_bins : 0u,
                            bp ? bp->max_leaf : 0u, P, !device_build, &cdf);
        if (msg.empty()) {
            bool unsupported = false;
            msg = prepare_sdf(desc->sdf_clusters, desc->sdf_cluster_count, P, &unsupported);
            if (unsupported) return fail(DMT_ERR_UNSUPPORTED, msg);
        }
    } catch (const std::bad_alloc&) {
        return fail(DMT_ERR_OUT_OF_MEMORY, "host allocation during scene build");
    }
    if (!msg.empty()) return fail(DMT_ERR_INVALID_ARG, msg);
    if (desc->mesh_uvs && P.tri_count)
        for (size_t i = 0; i < (size_t)desc->vertex_count * 2u; ++i)
            if (!std::isfinite(desc->mesh_uvs[i]))
                return fail(DMT_ERR_INVALID_ARG, "mesh UV of vertex " + std::to_string(i / 2) + " is not finite");
    auto* s = new (std::nothrow) dmt_scene();
    if (!s) return fail(DMT_ERR_OUT_OF_MEMORY, "host allocation");
    s->dev = dev;
    dev->scenes++;
    SceneView& V = s->view;
    std::memset(&V, 0, sizeof(V));
    auto bail = [&](dmt_status code) {
        dmt_scene_destroy(s);
        return code;
    };
    const float4* f4 = nullptr;
    if ((st = upload(s, (const float4*)P.prims.data(), (size_t)P.prim_count * 7u, &V.prims))) return bail(st);
    s->prims_dev = const_cast<float4*>(V.prims);   // allocated by upload(); written by dmt_scene_set_materials
    try {
        s->prims_host = P.prims;
    } catch (const std::bad_alloc&) {
        return bail(fail(DMT_ERR_OUT_OF_MEMORY, "host allocation"));
    }
    V.prim_count = P.prim_count;
    V.plane_count = (uint32_t)P.plane_ids.size();
    if ((st = upload(s, (const float4*)P.planes.data(), P.plane_ids.size(), &V.planes))) return bail(st);
    if ((st = upload(s, P.plane_ids.data(), P.plane_ids.size(), &V.plane_ids))) return bail(st);
    if ((st = upload(s, (const float4*)P.sphere_bvh.nodes.data(), P.sphere_bvh.nodes.size() * 8u, &f4))) return bail(st);
    V.sph_nodes = f4;
    if ((st = upload(s, (const float4*)P.spheres.data(), P.sphere_ids.size(), &V.spheres, kLeafPad))) return bail(st);
    if ((st = upload(s, P.sphere_ids.data(), P.sphere_ids.size(), &V.sphere_ids))) return bail(st);
    if ((st = upload(s, (const float4*)P.sphere_prev.data(), P.sphere_ids.size(), &V.sph_prev))) return bail(st);
    uint32_t tri_nodes = (uint32_t)P.tri_bvh.nodes.size(), tri_depth = P.tri_bvh.depth;
    std::vector<uint32_t> level_start;   // device build: first node of each level, node count last
    if (device_build && P.tri_count) {
        float* pos = nullptr;
        uint32_t* idx = nullptr;
        hipError_t e = dalloc(&pos, (size_t)desc->vertex_count * 3u);
        if (e == hipSuccess) s->buffers.push_back(pos);
        if (e == hipSuccess) e = dalloc(&idx, (size_t)P.tri_count * 3u);
        if (e == hipSuccess)
            e = hipMemcpy(pos, desc->positions, (size_t)desc->vertex_count * 12u, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(idx, desc->indices, (size_t)P.tri_count * 12u, hipMemcpyHostToDevice);
        LbvhOutput L;
        if (e == hipSuccess) {
            LbvhInput in{};
            in.positions = pos;
            in.indices = idx;
            in.tri_count = P.tri_count;
            for (int a = 0; a < 3; ++a) {
                in.cmin[a] = P.tri_cmin[a];
                in.cmax[a] = P.tri_cmax[a];
            }
            in.max_leaf = (bp && bp->max_leaf) ? bp->max_leaf : 4u;
            e = build_lbvh(in, kLeafPad, L, dev->stream);
        }
        if (idx) (void)hipFree(idx);
        if (e != hipSuccess) return bail(hip_fail(e, "device BVH build"));
        for (void* p : {(void*)L.nodes, (void*)L.tris, (void*)L.leaf_idx}) s->buffers.push_back(p);
        V.tri_nodes = L.nodes;
        V.tris = L.tris;
        s->leaf_idx = L.leaf_idx;
        s->positions_dev = pos;   // the device copy a later refit overwrites
        s->scene_bytes += (size_t)L.node_count * 128u + ((size_t)P.tri_count * 3u + kLeafPad) * 16u;
        tri_nodes = L.node_count;
        tri_depth = (uint32_t)L.level_start.size() - 1u;
        level_start = std::move(L.level_start);
    } else {
        if ((st = upload(s, (const float4*)P.tri_bvh.nodes.data(), P.tri_bvh.nodes.size() * 8u, &f4))) return bail(st);
        V.tri_nodes = f4;
        if ((st = upload(s, (const float4*)P.tris.data(), (size_t)P.tri_count * 3u, &V.tris, kLeafPad))) return bail(st);
    }
    if ((st = upload(s, (const float4*)P.lights.data(), (size_t)P.light_count * 5u, &V.lights))) return bail(st);
    V.light_count = P.light_count;
    V.ltree_count = (uint32_t)(P.light_tree.size() / 16u);
    if ((st = upload(s, (const float4*)P.light_tree.data(), (size_t)V.ltree_count * 4u, &V.ltree))) return bail(st);
    if (P.env_present) {
        if ((st = upload(s, (const float4*)P.env_rgba.data(), (size_t)P.env_w * P.env_h, &V.env))) return bail(st);
        if ((st = upload(s, P.env_marginal.data(), P.env_marginal.size(), &V.env_marg))) return bail(st);
        if ((st = upload(s, P.env_conditional.data(), P.env_conditional.size(), &V.env_cond))) return bail(st);
        V.env_w = P.env_w;
        V.env_h = P.env_h;
        V.env_total = P.env_total;
        V.env_intensity = P.env_intensity;
        V.env_present = 1u;
    }
    if (P.sdf_count) {
        if ((st = upload(s, (const float4*)P.sdf.data(), (size_t)P.sdf_count * 28u, &V.sdf))) return bail(st);
        if ((st = upload(s, (const float4*)P.sdf_bvh.nodes.data(), P.sdf_bvh.nodes.size() * 8u, &f4))) return bail(st);
        V.sdf_nodes = f4;
        if ((st = upload(s, (const float4*)P.sdf_leaf.data(), P.sdf_count, &V.sdf_leaf, kLeafPad))) return bail(st);
        V.sdf_count = P.sdf_count;
    }
    // queries without a frame (dmt_trace_closest) trace the clusters with the
    // engine's defaults, r_sdf_max_iters 128 and r_sdf_epsilon 1e-4
    V.sdf_iters = 128u;
    V.sdf_eps = 1e-4f;
    if (P.tri_count && device_build) {
        // the device build allocates its nodes level by level: the refit's
        // lists are those index ranges, deepest first
        std::vector<uint32_t> order;
        order.reserve(tri_nodes);
        for (size_t l = level_start.size() - 1u; l-- > 0;) {
            s->refit_level_start.push_back((uint32_t)order.size());
            for (uint32_t n = level_start[l]; n < level_start[l + 1u]; ++n) order.push_back(n);
        }
        s->refit_level_start.push_back((uint32_t)order.size());
        hipError_t e = dalloc(&s->refit_nodes, order.size());
        if (e == hipSuccess) s->buffers.push_back(s->refit_nodes);
        if (e == hipSuccess) e = hipMemcpy(s->refit_nodes, order.data(), order.size() * 4u, hipMemcpyHostToDevice);
        if (e != hipSuccess) return bail(hip_fail(e, "refit tables"));
        s->mesh_vertices = desc->vertex_count;
        s->mesh_tris = P.tri_count;
    } else if (P.tri_count) {
        // what a refit needs: the vertex indices of every leaf slot, and the
        // internal nodes grouped by level, deepest first
        std::vector<uint32_t> idx((size_t)P.tri_count * 3u);
        for (uint32_t k = 0; k < P.tri_count; ++k)
            for (uint32_t a = 0; a < 3u; ++a) idx[3u * k + a] = desc->indices[3u * P.tri_bvh.order[k] + a];
        std::vector<std::vector<uint32_t>> levels(1, std::vector<uint32_t>{0u});
        while (true) {
            std::vector<uint32_t> next;
            for (uint32_t n : levels.back())
                for (uint32_t c : P.tri_bvh.nodes[n].child)
                    if (c != kEmptyChild && !(c & kLeafBit)) next.push_back(c);
            if (next.empty()) break;
            levels.push_back(std::move(next));
        }
        std::vector<uint32_t> order;
        order.reserve(P.tri_bvh.nodes.size());
        for (size_t l = levels.size(); l-- > 0;) {
            s->refit_level_start.push_back((uint32_t)order.size());
            order.insert(order.end(), levels[l].begin(), levels[l].end());
        }
        s->refit_level_start.push_back((uint32_t)order.size());
        hipError_t e = dalloc(&s->leaf_idx, idx.size());
        if (e == hipSuccess) s->buffers.push_back(s->leaf_idx);
        if (e == hipSuccess) e = hipMemcpy(s->leaf_idx, idx.data(), idx.size() * 4u, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = dalloc(&s->refit_nodes, order.size());
        if (e == hipSuccess) s->buffers.push_back(s->refit_nodes);
        if (e == hipSuccess) e = hipMemcpy(s->refit_nodes, order.data(), order.size() * 4u, hipMemcpyHostToDevice);
        if (e != hipSuccess) return bail(hip_fail(e, "refit tables"));
        s->mesh_vertices = desc->vertex_count;
        s->mesh_tris = P.tri_count;
    }
    for (uint32_t i = 0; i < P.prim_count && !s->has_water; ++i) s->has_water = P.prims[(size_t)i * 28u + 9u] == 3.0f;
    for (uint32_t i = 0; i < P.sdf_count && !s->has_water; ++i) {
        uint32_t m;
        std::memcpy(&m, &P.sdf[(size_t)i * kSdfClusterFloats + 7u], 4);
        s->has_water = m == 3u;
    }
    if (desc->mesh_uvs && P.tri_count) {
        // per-vertex UVs, read through the vertex indices of each leaf slot
        if ((st = upload(s, (const float2*)desc->mesh_uvs, desc->vertex_count, &V.mesh_uv))) return bail(st);
        V.mesh_tri_vtx = s->leaf_idx;
    }
    s->bvh_depth = std::max(std::max(tri_depth, P.sphere_bvh.depth), P.sdf_bvh.depth);
    s->bvh_nodes = (uint32_t)(tri_nodes + P.sphere_bvh.nodes.size() + P.sdf_bvh.nodes.size());
    hipError_t e = hipMalloc((void**)&s->count_stats, sizeof(DeviceStats));
    if (e != hipSuccess) return bail(hip_fail(e, "hipMalloc(stats)"));
    s->buffers.push_back(s->count_stats);
    e = hipMalloc((void**)&s->stats, sizeof(DeviceStats));
    if (e != hipSuccess) return bail(hip_fail(e, "hipMalloc(stats)"));
    s->buffers.push_back(s->stats);
    e = hipMemset(s->stats, 0, sizeof(DeviceStats));
    for (uint32_t k = 0; k < kFrameSets && e == hipSuccess; ++k) {
        e = hipEventCreateWithFlags(&s->acc_done[k], hipEventDisableTiming);
        for (uint32_t j = 0; j < kPollSlots && e == hipSuccess; ++j)
            e = hipEventCreateWithFlags(&s->poll[k][j], hipEventDisableTiming);
    }
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&s->host_count, kFrameSets * kPollSlots * sizeof(uint32_t), hipHostMallocDefault);
    if (e != hipSuccess) return bail(hip_fail(e, "scene events"));
    *out = s;
    return DMT_OK;
}

}  // extern "C"

namespace {

void scene_free(dmt_scene* s) {
    if (s->dev) {
        (void)hipSetDevice(s->dev->id);
        (void)sync_all(s->dev);
    }
    free_accum(s);
    for (hipEvent_t e : s->acc_done)
        if (e) (void)hipEventDestroy(e);
    for (auto& row : s->poll)
        for (hipEvent_t e : row)
            if (e) (void)hipEventDestroy(e);
    if (s->host_count) (void)hipHostFree(s->host_count);
    if (s->adj_spill) (void)hipFree(s->adj_spill);
    if (s->adj_rec) (void)hipFree(s->adj_rec);
    if (s->adj_rec_count) (void)hipFree(s->adj_rec_count);
    if (s->moon) (void)hipFree(s->moon);
    if (s->atlas) (void)hipFree(s->atlas);
    for (void* p : s->buffers) (void)hipFree(p);
    for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
    dmt_device* d = s->dev;
    delete s;
    if (d && --d->scenes == 0 && d->destroy_requested) device_free(d);
}

}  // namespace

extern "C" {

void dmt_scene_destroy(dmt_scene* s) {
    if (!s) return;
    if (s->batches > 0) {
        s->destroy_requested = true;
        return;
    }
    scene_free(s);
}

namespace {

// Renders frame_count consecutive frames, each on the next wavefront set and
// its stream: frames stay in flight while the accumulation passes are ordered
// frame after frame by events.
dmt_status render_impl(dmt_scene* s, const dmt_frame_params* f, uint32_t frame_count) {
    if (!s) return fail(DMT_ERR_INVALID_ARG, "scene is null");
    dmt_status st = check_device(s->dev);
    if (st) return st;
    if ((st = check_frame_for(s, f))) return st;
    if (frame_count == 0) return fail(DMT_ERR_INVALID_ARG, "frame_count must be >= 1");
    bool resized = false;
    if ((st = ensure_wavefront(s, f->width, f->height, &resized))) return st;
    if (sky_lights(s, f) && (st = ensure_sky_queues(s))) return st;
    if (f->max_bounces == 0) return DMT_OK;
    for (uint32_t k = 0; k < frame_count; ++k) {
        FrameArgs F = frame_args(f);
        F.frame_index = f->frame_index + k;
        F.reset_accum = (k == 0u && (f->reset_accum || resized)) ? 1u : 0u;
        const int set = s->next_set;
        hipStream_t stream = set_stream(s, set);
        LaunchOptions o;
        o.poll_events = s->poll[set];
        o.accum_after = s->last_set >= 0 ? s->acc_done[s->last_set] : nullptr;
        if (s->profiling) {
            const size_t bound = 2u * (size_t)f->spp * f->max_bounces;
            while (s->ev_pool.size() < s->ev_used + bound) {
                hipEvent_t e;
                hipError_t he = hipEventCreate(&e);
                if (he != hipSuccess) return hip_fail(he, "hipEventCreate");
                s->ev_pool.push_back(e);
            }
            if (s->ev_kind.size() < s->ev_pool.size() / 2) s->ev_kind.resize(s->ev_pool.size() / 2);
            o.extend_events = s->ev_pool.data() + s->ev_used;
            o.event_kinds = s->ev_kind.data() + s->ev_used / 2;
            o.max_events = (uint32_t)bound;
        }
        o.tail_after = s->tail_after;
        const FeatureArgs X = feature_args(s, f);
        if (uses_features(s, f)) o.features = &X;
        o.tail_blocks = s->dev->cus * (o.features ? tail_waves(X.ext) : kTailBlocksPerCu);
        o.extend_blocks = s->dev->cus * kExtendBlocksPerCu;
        hipError_t e = launch_frame(frame_view(s, f), F, s->wf[set], s->accum, stream, o);
        s->ev_used += o.events_used;
        if (e == hipSuccess) e = hipEventRecord(s->acc_done[set], stream);
        if (e != hipSuccess) return mark_lost_if_fatal(s->dev, e, "launch_frame");
        s->last_set = set;
        s->next_set = (set + 1) % (int)kFrameSets;
        s->frames++;
    }
    return DMT_OK;
}

}  // namespace

dmt_status dmt_render(dmt_scene* s, const dmt_frame_params* f) { return render_impl(s, f, 1u); }

dmt_status dmt_render_frames(dmt_scene* s, const dmt_frame_params* f, uint32_t frame_count) {
    return render_impl(s, f, frame_count);
}

dmt_status dmt_readback_accum(dmt_scene* s, float* rgba, size_t count) {
    if (!s || !rgba) return fail(DMT_ERR_INVALID_ARG, "null pointer");
    dmt_status st = check_device(s->dev);
    if (st) return st;
    if (!s->accum) return fail(DMT_ERR_INVALID_ARG, "nothing rendered yet");
    if (count != (size_t)s->acc_w * s->acc_h * 4u) return fail(DMT_ERR_INVALID_ARG, "size mismatch (full readback only)");
    hipError_t e = sync_all(s->dev);
    if (e == hipSuccess) e = hipMemcpyAsync(rgba, s->accum, count * 4u, hipMemcpyDeviceToHost, s->dev->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->dev->stream);
    if (e != hipSuccess) return mark_lost_if_fatal(s->dev, e, "readback");
    return DMT_OK;
}

dmt_status dmt_accum_device_ptr(dmt_scene* s, void** out) {
    if (!s || !out) return fail(DMT_ERR_INVALID_ARG, "null pointer");
    *out = s->accum;
    return s->accum ? DMT_OK : fail(DMT_ERR_INVALID_ARG, "nothing rendered yet");
}

dmt_status dmt_scene_bind_accum(dmt_scene* s, void* d_rgba, uint32_t w, uint32_t h) {
    if (!s) return fail(DMT_ERR_INVALID_ARG, "scene is null");
    dmt_status st = check_device(s->dev);
    if (st) return st;
    if (s->batches > 0) return fail(DMT_ERR_INVALID_ARG, "the scene belongs to a batch");
    if (d_rgba) {
        if (w == 0 || h == 0) return fail(DMT_ERR_INVALID_ARG, "width and height must be >= 1");
        if (w > 65535u || h > 65535u) return fail(DMT_ERR_LIMIT, "width and height must be <= 65535");
        if (reinterpret_cast<uintptr_t>(d_rgba) % 16u) return fail(DMT_ERR_INVALID_ARG, "d_rgba must be 16-byte aligned");
        hipPointerAttribute_t attr{};
        hipError_t pe = hipPointerGetAttributes(&attr, d_rgba);
        if (pe != hipSuccess) {
            (void)hipGetLastError();
            return fail(DMT_ERR_INVALID_ARG, "d_rgba is not a device allocation");
        }
        if (attr.type != hipMemoryTypeDevice || attr.device != s->dev->id)
            return fail(DMT_ERR_INVALID_ARG, "d_rgba is not device memory of the scene's device");
    }
    // frames in flight write the current buffer
    hipError_t e = sync_all(s->dev);
    if (e != hipSuccess) return mark_lost_if_fatal(s->dev, e, "synchronize");
    if (!d_rgba) {
        if (s->accum_external) free_accum(s);
        return DMT_OK;
    }
    if (s->accum && s->acc_w == w && s->acc_h == h) {
        // same resolution: the wavefront sets stay
        if (!s->accum_external) (void)hipFree(s->accum);
    } else {
        free_accum(s);
        s->acc_w = w;
        s->acc_h = h;
    }
    s->accum = static_cast<float4*>(d_rgba);
    s->accum_external = true;
    return DMT_OK;
}

dmt_status dmt_tonemap(dmt_scene* s, const dmt_tonemap_params* p, void* d_out) {
    if (!s || !p || !d_out) return fail(DMT_ERR_INVALID_ARG, "null pointer");
    dmt_status st = check_device(s->dev);
    if (st) return st;
    if (!s->accum) return fail(DMT_ERR_INVALID_ARG, "nothing rendered yet");
    if (reinterpret_cast<uintptr_t>(d_out) % 16u) return fail(DMT_ERR_INVALID_ARG, "d_out_rgba must be 16-byte aligned");
    if (p->mode > DMT_TONEMAP_HDR) return fail(DMT_ERR_INVALID_ARG, "unknown tonemap mode");
    const float exposure = (std::isfinite(p->exposure) && p->exposure > 0.0f) ? p->exposure : 1.0f;
    hipStream_t stream = s->dev->stream;
    hipError_t e = hipSuccess;
    if (s->last_set >= 0) e = hipStreamWaitEvent(stream, s->acc_done[s->last_set], 0);
    if (e == hipSuccess)
        e = launch_tonemap(s->accum, s->acc_w * s->acc_h, exposure, p->op, p->mode, static_cast<float4*>(d_out),
                           stream);
    // the next frame's accumulation pass waits for this read
    if (e == hipSuccess && s->last_set >= 0) e = hipEventRecord(s->acc_done[s->last_set], stream);
    if (e != hipSuccess) return mark_lost_if_fatal(s->dev, e, "tonemap");
    return DMT_OK;
}

dmt_status dmt_scene_update_mesh(dmt_scene* s, const float* positions, uint32_t vertex_count) {
    if (!s || (!positions && vertex_count)) return fail(DMT_ERR_INVALID_ARG, "null pointer");
    dmt_status st = check_device(s->dev);
    if (st) return st;
    if (s->mesh_tris == 0) return fail(DMT_ERR_INVALID_ARG, "the scene has no mesh");
    if (vertex_count != s->mesh_vertices)
        return fail(DMT_ERR_INVALID_ARG, "vertex_count must be the mesh's " + std::to_string(s->mesh_vertices));
    const size_t nf = (size_t)vertex_count * 3u;
    for (size_t i = 0; i < nf; ++i)
        if (!std::isfinite(positions[i])) return fail(DMT_ERR_INVALID_ARG, "positions contain a non-finite value");
    hipError_t e = hipSuccess;
    if (!s->positions_dev) {
        e = dalloc(&s->positions_dev, nf);
        if (e != hipSuccess) {
            s->positions_dev = nullptr;
            return hip_fail(e, "hipMalloc(positions)");
        }
        s->buffers.push_back(s->positions_dev);
    }
    // frames in flight read the triangles and the tree
    e = sync_all(s->dev);
    if (e == hipSuccess) e = hipMemcpy(s->positions_dev, positions, nf * 4u, hipMemcpyHostToDevice);
    hipStream_t stream = s->dev->stream;
    if (e == hipSuccess)
        e = launch_refit_tris(s->positions_dev, s->leaf_idx, s->mesh_tris, const_cast<float4*>(s->view.tris), stream);
    for (size_t l = 0; e == hipSuccess && l + 1 < s->refit_level_start.size(); ++l) {
        const uint32_t b = s->refit_level_start[l], n = s->refit_level_start[l + 1] - b;
        e = launch_refit_nodes(const_cast<float4*>(s->view.tri_nodes), s->refit_nodes + b, n, s->positions_dev,
                               s->leaf_idx, stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return mark_lost_if_fatal(s->dev, e, "mesh refit");
    return DMT_OK;
}

dmt_status dmt_render_gbuffer(dmt_scene* s, const dmt_frame_params* f, const dmt_gbuffer_params* p) {
    if (!s || !p) return fail(DMT_ERR_INVALID_ARG, "null pointer");
    dmt_status st = check_device(s->dev);
    if (st) return st;
    if ((st = check_frame_for(s, f))) return st;
    GbufferArgs A{};
    A.F = frame_args(f);
    std::memcpy(A.vp_curr, p->curr_view_proj, sizeof(A.vp_curr));
    std::memcpy(A.vp_prev, p->prev_view_proj, sizeof(A.vp_prev));
    A.stride = p->pixel_stride;
    float* const outs[kGbOutputs] = {p->depth,           p->motion,    p->normal,
                                     p->albedo,          p->specular_albedo, p->roughness,
                                     p->specular_hit_distance, p->cloud_transmittance};
    static constexpr uint32_t kFloats[kGbOutputs] = {1u, 2u, 4u, 4u, 4u, 1u, 1u, 1u};
    bool any = false;
    for (uint32_t k = 0; k < kGbOutputs; ++k) {
        A.out[k] = outs[k];
        if (!outs[k]) continue;
        any = true;
        if (A.stride < kFloats[k])
            return fail(DMT_ERR_INVALID_ARG, "pixel_stride " + std::to_string(A.stride) + " is below the " +
                                                 std::to_string(kFloats[k]) + " floats output " + std::to_string(k) +
                                                 " writes");
        if (reinterpret_cast<uintptr_t>(outs[k]) % 4u) return fail(DMT_ERR_INVALID_ARG, "outputs must be float-aligned");
    }
    if (!any) return DMT_OK;
    bool resized = false;
    if (s->wf[0].capacity == 0u && (st = ensure_wavefront(s, f->width, f->height, &resized))) return st;
    // on the device stream, as frame set 0, whose traversal spill area it shares
    hipError_t e = launch_gbuffer(frame_view(s, f), A, s->wf[0], s->dev->stream);
    if (e != hipSuccess) return mark_lost_if_fatal(s->dev, e, "gbuffer");
    return DMT_OK;
}

dmt_status dmt_trace_closest(dmt_scene* s, const float* rays, uint32_t n, dmt_hit* out) {
    static_assert(sizeof(dmt_hit) == sizeof(HitOut), "dmt_hit layout");
    if (!s || (n && (!rays || !out))) return fail(DMT_ERR_INVALID_ARG, "null pointer");
    dmt_status st = check_device(s->dev);
    if (st) return st;
    if (n == 0) return DMT_OK;
    bool resized = false;
    if (s->wf[0].capacity == 0u &&
        (st = ensure_wavefront(s, s->accum ? s->acc_w : 1u, s->accum ? s->acc_h : 1u, &resized)))
        return st;
    float4* d_rays = nullptr;
    HitOut* d_out = nullptr;
    hipError_t e = hipMalloc((void**)&d_rays, (size_t)n * 32u);
    if (e == hipSuccess) e = hipMalloc((void**)&d_out, (size_t)n * sizeof(HitOut));
    if (e == hipSuccess) e = hipMemcpyAsync(d_rays, rays, (size_t)n * 32u, hipMemcpyHostToDevice, s->dev->stream);
    if (e == hipSuccess) e = launch_trace_query(s->view, d_rays, n, d_out, s->wf[0], s->dev->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, d_out, (size_t)n * sizeof(HitOut), hipMemcpyDeviceToHost, s->dev->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->dev->stream);
    (void)hipFree(d_rays);
    (void)hipFree(d_out);
    if (e != hipSuccess) return mark_lost_if_fatal(s->dev, e, "trace query");
    return DMT_OK;
}

dmt_status dmt_scene_set_moon_map(dmt_scene* s, const float* rgba, uint32_t w, uint32_t h) {
    if (!s) return fail(DMT_ERR_INVALID_ARG, "scene is null");
    dmt_status st = check_device(s->dev);
    if (st) return st;
    if (s->batches > 0) return fail(DMT_ERR_INVALID_ARG, "the scene belongs to a batch");
    if (rgba) {
        if (w == 0 || h == 0) return fail(DMT_ERR_INVALID_ARG, "moon texture width and height must be >= 1");
        if ((uint64_t)w * h > (1ull << 24)) return fail(DMT_ERR_LIMIT, "moon texture larger than 2^24 texels");
        for (size_t i = 0; i < (size_t)w * h * 4u; ++i)
            if (!std::isfinite(rgba[i])) return fail(DMT_ERR_INVALID_ARG, "moon texture has a non-finite value");
    }
    hipError_t e = sync_all(s->dev);   // frames in flight read the old texture
    if (e != hipSuccess) return mark_lost_if_fatal(s->dev, e, "synchronize");
    if (s->moon) (void)hipFree(s->moon);
    s->moon = nullptr;
    s->view.moon = nullptr;
    s->view.moon_w = s->view.moon_h = 0u;
    if (!rgba) return DMT_OK;
    const size_t bytes = (size_t)w * h * 16u;
    e = hipMalloc((void**)&s->moon, bytes);
    if (e == hipSuccess) e = hipMemcpy(s->moon, rgba, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (s->moon) (void)hipFree(s->moon);
        s->moon = nullptr;
        return hip_fail(e, "moon texture upload");
    }
    s->view.moon = s->moon;
    s->view.moon_w = w;
    s->view.moon_h = h;
    return DMT_OK;
}

dmt_status dmt_scene_set_pbr_atlas(dmt_scene* s, const uint8_t* rgba8, uint32_t w, uint32_t h) {
    if (!s) return fail(DMT_ERR_INVALID_ARG, "scene is null");
    dmt_status st = check_device(s->dev);
    if (st) return st;
    if (s->batches > 0) return fail(DMT_ERR_INVALID_ARG, "the scene belongs to a batch");
    if (rgba8) {
        if (w != kPbrAtlasW) return fail(DMT_ERR_INVALID_ARG, "PBR atlas width must be 384");
        if (h == 0 || h % kPbrTile != 0u) return fail(DMT_ERR_INVALID_ARG, "PBR atlas height must be a multiple of 256");
        if (h / kPbrTile > 4096u) return fail(DMT_ERR_LIMIT, "PBR atlas larger than 4096 tiles");
    }
    hipError_t e = sync_all(s->dev);   // frames in flight read the old atlas
    if (e != hipSuccess) return mark_lost_if_fatal(s->dev, e, "synchronize");
    if (s->atlas) (void)hipFree(s->atlas);
    s->atlas = nullptr;
    s->view.atlas = nullptr;
    s->view.atlas_tiles = 0u;
    s->prims_textured = false;
    if (!rgba8) return DMT_OK;
    const size_t bytes = (size_t)w * h * 4u;
    e = hipMalloc((void**)&s->atlas, bytes);
    if (e == hipSuccess) e = hipMemcpy(s->atlas, rgba8, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (s->atlas) (void)hipFree(s->atlas);
        s->atlas = nullptr;
        return hip_fail(e, "PBR atlas upload");
    }
    s->view.atlas = s->atlas;
    s->view.atlas_tiles = h / kPbrTile;
    const size_t n = s->prims_host.size() / 28u;
    for (size_t i = 0; i < n && !s->prims_textured; ++i)
        for (int c = 24; c < 28; ++c) {
            uint32_t t;
            std::memcpy(&t, &s->prims_host[i * 28u + c], 4);
            if (t < s->view.atlas_tiles) s->prims_textured = true;
        }
    return DMT_OK;
}

dmt_status dmt_trace_closest_noalloc(dmt_scene* s, const float* d_rays, uint32_t n, dmt_hit* d_out) {
    static_assert(sizeof(dmt_hit) == sizeof(HitOut), "dmt_hit layout");
    if (!s || (n && (!d_rays || !d_out))) return fail(DMT_ERR_INVALID_ARG, "null pointer");
    if (((uintptr_t)d_rays & 15u) != 0u || ((uintptr_t)d_out & 3u) != 0u)
        return fail(DMT_ERR_INVALID_ARG, "d_rays must be 16-byte and d_out 4-byte aligned");
    dmt_status st = check_device(s->dev);
    if (st) return st;
    if (!s->dev->noalloc_scratch) return fail(DMT_ERR_DEVICE, s->dev->noalloc_error);
    // the deepest traversal stack of the scene's trees: 3 entries per level
    // (trace_closest pushes at most three children per node), the first
    // kLdsDepth in LDS
    const uint32_t need = 3u * s->bvh_depth + 4u;
    const uint32_t spill = need > kLdsDepth ? need - kLdsDepth : 0u;
    if (spill > DMT_NOALLOC_SPILL_DEPTH)
        return fail(DMT_ERR_LIMIT, "scene BVH depth " + std::to_string(s->bvh_depth) + " needs " +
                                       std::to_string(spill) + " spilled stack entries per thread; the static scratch "
                                       "holds " + std::to_string(DMT_NOALLOC_SPILL_DEPTH) + " (BVH depth <= " +
                                       std::to_string((DMT_NOALLOC_SPILL_DEPTH + kLdsDepth - 4u) / 3u) + ")");
    if (n == 0) return DMT_OK;
    const hipError_t e = launch_trace_query_spill(s->view, (const float4*)d_rays, n, (HitOut*)d_out,
                                                  s->dev->noalloc_scratch, DMT_NOALLOC_SPILL_DEPTH, DMT_NOALLOC_BLOCKS,
                                                  s->dev->stream);
    if (e != hipSuccess) return mark_lost_if_fatal(s->dev, e, "trace query (no allocation)");
    return DMT_OK;
}

dmt_status dmt_scene_stats(dmt_scene* s, dmt_stats* out) {
    if (!s || !out) return fail(DMT_ERR_INVALID_ARG, "null pointer");
    dmt_status st = check_device(s->dev);
    if (st) return st;
    std::memset(out, 0, sizeof(*out));
    if ((st = harvest_events(s))) return st;
    {
        DeviceStats h{};
        hipError_t e = hipMemcpyAsync(&h, s->stats, sizeof(h), hipMemcpyDeviceToHost, s->dev->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->dev->stream);
        if (e != hipSuccess) return mark_lost_if_fatal(s->dev, e, "stats");
        out->primary_rays = h.primary;
        out->extension_rays = h.extension;
        out->shadow_rays = h.shadow;
    }
    out->frames = s->frames;
    out->extend_launches = s->extend_launches;
    out->extend_ms = s->extend_ms;
    out->tail_launches = s->tail_launches;
    out->tail_ms = s->tail_ms;
    out->scene_bytes = s->scene_bytes;
    out->bvh_nodes = s->bvh_nodes;
    out->bvh_depth = s->bvh_depth;
    return DMT_OK;
}

dmt_status dmt_scene_reset_stats(dmt_scene* s) {
    if (!s) return fail(DMT_ERR_INVALID_ARG, "scene is null");
    dmt_status st = check_device(s->dev);
    if (st) return st;
    if ((st = harvest_events(s))) return st;
    s->extend_ms = 0.0;
    s->extend_launches = 0;
    s->tail_ms = 0.0;
    s->tail_launches = 0;
    s->frames = 0;
    {
        hipError_t e = hipMemsetAsync(s->stats, 0, sizeof(DeviceStats), s->dev->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->dev->stream);
        if (e != hipSuccess) return hip_fail(e, "hipMemsetAsync");
    }
    return DMT_OK;
}

dmt_status dmt_scene_set_schedule(dmt_scene* s, uint32_t wavefront_iterations) {
    if (!s) return fail(DMT_ERR_INVALID_ARG, "scene is null");
    s->tail_after = wavefront_iterations;
    return DMT_OK;
}

dmt_status dmt_scene_get_schedule(dmt_scene* s, uint32_t* wavefront_iterations) {
    if (!s || !wavefront_iterations) return fail(DMT_ERR_INVALID_ARG, "null pointer");
    *wavefront_iterations = s->tail_after;
    return DMT_OK;
}

dmt_status dmt_scene_set_profiling(dmt_scene* s, int32_t enabled) {
    if (!s) return fail(DMT_ERR_INVALID_ARG, "scene is null");
    s->profiling = enabled != 0;
    return DMT_OK;
}

dmt_status dmt_scene_measure_traffic(dmt_scene* s, const dmt_frame_params* f, dmt_stats* out) {
    if (!s || !out) return fail(DMT_ERR_INVALID_ARG, "null pointer");
    dmt_status st = check_device(s->dev);
    if (st) return st;
    if ((st = check_frame_for(s, f))) return st;
    bool resized = false;
    if ((st = ensure_wavefront(s, f->width, f->height, &resized))) return st;
    if (sky_lights(s, f) && (st = ensure_sky_queues(s))) return st;
    // Render into a scratch accumulation buffer so the scene's image and
    // counters are left as they were.
    float4* scratch = nullptr;
    hipError_t e = hipMalloc((void**)&scratch, (size_t)f->width * f->height * 16u);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(scratch)");
    e = sync_all(s->dev);
    if (e != hipSuccess) { (void)hipFree(scratch); return mark_lost_if_fatal(s->dev, e, "synchronize"); }
    Wavefront W = s->wf[0];
    W.stats = s->count_stats;
    e = hipMemsetAsync(s->count_stats, 0, sizeof(DeviceStats), s->dev->stream);
    FrameArgs F = frame_args(f);
    F.reset_accum = 1u;
    LaunchOptions o;
    o.counting = true;
    o.tail_after = s->tail_after;
    const FeatureArgs X = feature_args(s, f);
    if (uses_features(s, f)) o.features = &X;
    o.tail_blocks = s->dev->cus * (o.features ? tail_waves(X.ext) : kTailBlocksPerCu);
    o.extend_blocks = s->dev->cus * kExtendBlocksPerCu;
    if (e == hipSuccess) e = launch_frame(frame_view(s, f), F, W, scratch, s->dev->stream, o);
    DeviceStats h{};
    if (e == hipSuccess) e = hipMemcpyAsync(&h, s->count_stats, sizeof(h), hipMemcpyDeviceToHost, s->dev->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->dev->stream);
    (void)hipFree(scratch);
    if (e != hipSuccess) return mark_lost_if_fatal(s->dev, e, "measure_traffic");
    std::memset(out, 0, sizeof(*out));
    out->primary_rays = h.primary;
    out->extension_rays = h.extension;
    out->shadow_rays = h.shadow;
    out->node_visits = h.nodes + h.ext_nodes;
    out->tri_tests = h.tris + h.ext_tris;
    out->sphere_tests = h.spheres + h.ext_spheres;
    out->ext_node_visits = h.ext_nodes;
    out->ext_tri_tests = h.ext_tris;
    out->ext_sphere_tests = h.ext_spheres;
    out->ext_rays = h.ext_rays;
    out->tail_node_visits = h.tail_nodes;
    out->tail_tri_tests = h.tail_tris;
    out->tail_sphere_tests = h.tail_spheres;
    out->node_visits += h.tail_nodes;
    out->tri_tests += h.tail_tris;
    out->sphere_tests += h.tail_spheres;
    out->frames = 1;
    out->scene_bytes = s->scene_bytes;
    out->bvh_nodes = s->bvh_nodes;
    out->bvh_depth = s->bvh_depth;
    return DMT_OK;
}

// ---------------------------------------------------------------------------
// Material parameters and the adjoint.

dmt_status dmt_scene_get_materials(dmt_scene* s, float* out, size_t float_count) {
    if (!s || (!out && float_count)) return fail(DMT_ERR_INVALID_ARG, "null pointer");
    const size_t n = s->prims_host.size() / 28u;
    if (float_count != n * kMaterialFloats) return fail(DMT_ERR_INVALID_ARG, "float_count must be prim_count * 7");
    for (size_t i = 0; i < n; ++i) {
        const float* p = &s->prims_host[i * 28u];
        float* o = out + i * kMaterialFloats;
        o[0] = p[4]; o[1] = p[5]; o[2] = p[6]; o[3] = p[7];
        o[4] = p[16]; o[5] = p[17]; o[6] = p[18];
    }
    return DMT_OK;
}

dmt_status dmt_scene_set_materials(dmt_scene* s, const float* mats, size_t float_count) {
    if (!s || (!mats && float_count)) return fail(DMT_ERR_INVALID_ARG, "null pointer");
    dmt_status st = check_device(s->dev);
    if (st) return st;
    const size_t n = s->prims_host.size() / 28u;
    if (float_count != n * kMaterialFloats) return fail(DMT_ERR_INVALID_ARG, "float_count must be prim_count * 7");
    for (size_t i = 0; i < float_count; ++i)
        if (!std::isfinite(mats[i])) return fail(DMT_ERR_INVALID_ARG, "material " + std::to_string(i / 7) + " has a non-finite value");
    if (n == 0) return DMT_OK;
    // frames in flight read the current values
    hipError_t e = sync_all(s->dev);
    if (e != hipSuccess) return mark_lost_if_fatal(s->dev, e, "synchronize");
    for (size_t i = 0; i < n; ++i) {
        float* p = &s->prims_host[i * 28u];
        const float* m = mats + i * kMaterialFloats;
        p[4] = m[0]; p[5] = m[1]; p[6] = m[2]; p[7] = m[3];
        p[16] = m[4]; p[17] = m[5]; p[18] = m[6];
    }
    e = hipMemcpy(s->prims_dev, s->prims_host.data(), s->prims_host.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) return mark_lost_if_fatal(s->dev, e, "material upload");
    return DMT_OK;
}

dmt_status dmt_adjoint(dmt_scene* s, const dmt_frame_params* f, const void* d_image_grad, void* d_param_grad) {
    if (!s) return fail(DMT_ERR_INVALID_ARG, "scene is null");
    dmt_status st = check_device(s->dev);
    if (st) return st;
    if ((st = check_frame_for(s, f))) return st;
    if (f->max_bounces > kAdjMaxVerts) return fail(DMT_ERR_LIMIT, "the adjoint supports max_bounces <= 8");
    if (f->refract_bounces > kAdjMaxDielectric) return fail(DMT_ERR_LIMIT, "the adjoint supports refract_bounces <= 8");
    if (sky_lights(s, f))
        return fail(DMT_ERR_UNSUPPORTED, "the adjoint does not differentiate sky lights (sun, moon, HDRI clusters)");
    if (pbr_active(s, f))
        return fail(DMT_ERR_UNSUPPORTED, "the adjoint does not differentiate PBR texture maps");
    if (water_active(s, f)) return fail(DMT_ERR_UNSUPPORTED, "the adjoint does not differentiate water");
    const uint32_t prims = s->view.prim_count;
    if (prims && (!d_image_grad || !d_param_grad)) return fail(DMT_ERR_INVALID_ARG, "null device pointer");
    if (prims == 0) return DMT_OK;
    const uint32_t need = 3u * s->bvh_depth + 4u;
    const uint32_t spill_depth = need > kLdsDepth ? need - kLdsDepth : 0u;
    if (spill_depth && !s->adj_spill) {
        hipError_t e = hipMalloc((void**)&s->adj_spill, (size_t)kMaxGridBlocks * kBlock * spill_depth * 4u);
        if (e != hipSuccess) {
            s->adj_spill = nullptr;
            return hip_fail(e, "hipMalloc(adjoint spill)");
        }
    }
    AdjointArgs A{};
    A.S = frame_view(s, f);
    A.F = frame_args(f);
    A.X = feature_args(s, f);
    A.features = uses_features(s, f) ? 1u : 0u;
    A.dimage = (const float4*)d_image_grad;
    A.grads = (float*)d_param_grad;
    A.prim_count = prims;
    for (uint32_t i = 0; i < prims; ++i)
        if (s->prims_host[(size_t)i * 28u + 9u] == 1.0f && s->prims_host[(size_t)i * 28u + 8u] == 0.0f) A.metal_prims++;
    A.spill = s->adj_spill;
    A.spill_depth = spill_depth;
    if (A.metal_prims) {
        // at most one record per vertex of every sample; pixel chunks keep the
        // list within kAdjRecMaxBytes
        const uint64_t per_px = (uint64_t)f->spp * f->max_bounces;
        const uint64_t npix = (uint64_t)f->width * f->height;
        const uint64_t fit = kAdjRecMaxBytes / (kAdjRecFloat4 * sizeof(float4)) / per_px;
        A.chunk_px = (uint32_t)std::max<uint64_t>(1u, std::min<uint64_t>(npix, fit));
        A.rec_cap = (uint32_t)(A.chunk_px * per_px);
        const size_t bytes = (size_t)A.rec_cap * kAdjRecFloat4 * sizeof(float4);
        if (bytes > s->adj_rec_bytes) {
            if (s->adj_rec) (void)hipFree(s->adj_rec);
            s->adj_rec = nullptr;
            s->adj_rec_bytes = 0;
            hipError_t e = hipMalloc((void**)&s->adj_rec, bytes);
            if (e != hipSuccess) {
                s->adj_rec = nullptr;
                return hip_fail(e, "hipMalloc(adjoint records)");
            }
            s->adj_rec_bytes = bytes;
        }
        if (!s->adj_rec_count) {
            hipError_t e = hipMalloc((void**)&s->adj_rec_count, sizeof(uint32_t));
            if (e != hipSuccess) {
                s->adj_rec_count = nullptr;
                return hip_fail(e, "hipMalloc(adjoint record count)");
            }
        }
        A.rec = s->adj_rec;
        A.rec_count = s->adj_rec_count;
    }
    hipError_t e = launch_adjoint(A, s->dev->stream);
    if (e != hipSuccess) return mark_lost_if_fatal(s->dev, e, "adjoint");
    return DMT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched persistent form.

struct dmt_batch {
    dmt_device* dev = nullptr;
    std::vector<dmt_scene*> scenes;
    std::vector<dmt_frame_params> frames;
    dmt_batch_params params{};
    uint32_t cus = 0;
    uint32_t max_threads = 0;         // 4 workgroups per CU
    uint32_t spill_depth = 0;         // per thread, for the smallest LDS stack
    std::vector<uint32_t> order;      // scenes grouped by XCD queue
    uint32_t queue_begin[kXcds + 1] = {};
    BatchScene* d_scenes = nullptr;
    uint32_t* d_order = nullptr;
    uint32_t* d_work = nullptr;
    uint32_t* d_frame = nullptr;
    float4* d_rec = nullptr;
    uint32_t* d_spill = nullptr;
    DeviceStats* d_stats = nullptr;
    hipGraph_t graph = nullptr;       // one frame: reset, persistent kernel, frame counter
    hipGraphExec_t exec = nullptr;
    uint64_t frames_rendered = 0;
};

namespace {

constexpr uint32_t kBatchMaxScenes = 4096;
constexpr uint32_t kBatchMaxBlocksPerCu = 5;

dmt_status normalize_params(const dmt_batch_params* in, dmt_batch_params* out) {
    dmt_batch_params p{};
    if (in) p = *in;
    if (p.tile_w == 0) p.tile_w = 8;
    if (p.tiles_per_fetch == 0) p.tiles_per_fetch = 1;
    if (p.blocks_per_cu == 0) p.blocks_per_cu = 4;
    if (p.stack_lds == 0) p.stack_lds = 16;
    if (p.tile_w != 8 && p.tile_w != 16 && p.tile_w != 32 && p.tile_w != 64)
        return fail(DMT_ERR_INVALID_ARG, "tile_w must be 8, 16, 32 or 64");
    if (p.tiles_per_fetch > 16) return fail(DMT_ERR_INVALID_ARG, "tiles_per_fetch must be in 1..16");
    if (p.blocks_per_cu > kBatchMaxBlocksPerCu) return fail(DMT_ERR_INVALID_ARG, "blocks_per_cu must be in 1..5");
    if (p.stack_lds != 16 && p.stack_lds != 24 && p.stack_lds != 32)
        return fail(DMT_ERR_INVALID_ARG, "stack_lds must be 16, 24 or 32");
    if (p.plain_launches > 1) return fail(DMT_ERR_INVALID_ARG, "plain_launches must be 0 or 1");
    for (uint32_t r : p.reserved)
        if (r) return fail(DMT_ERR_INVALID_ARG, "reserved fields must be 0");
    *out = p;
    return DMT_OK;
}

uint32_t log2u(uint32_t v) {
    uint32_t r = 0;
    while ((1u << r) < v) ++r;
    return r;
}

// Scene table for tile width tile_w; accum[i] replaces scene i's buffer
// when given (autotuning renders into scratch buffers).
std::vector<BatchScene> batch_table(const dmt_batch* b, uint32_t tile_w, const std::vector<float4*>* accum) {
    std::vector<BatchScene> t(b->scenes.size());
    const uint32_t tile_h = 64u / tile_w;
    for (size_t i = 0; i < t.size(); ++i) {
        const dmt_frame_params& f = b->frames[i];
        BatchScene& e = t[i];
        std::memset(&e, 0, sizeof(e));
        e.S = frame_view(b->scenes[i], &f);
        e.F = frame_args(&f);
        e.X = feature_args(b->scenes[i], &f);
        e.accum = accum ? (*accum)[i] : b->scenes[i]->accum;
        e.tiles_x = (f.width + tile_w - 1u) / tile_w;
        const uint32_t tiles_y = (f.height + tile_h - 1u) / tile_h;
        e.tiles = f.max_bounces == 0u ? 0u : e.tiles_x * tiles_y;
    }
    return t;
}

BatchArgs batch_args(const dmt_batch* b, const dmt_batch_params& p, const BatchScene* table, DeviceStats* stats) {
    BatchArgs A{};
    A.scenes = table;
    A.order = b->d_order;
    for (uint32_t q = 0; q <= kXcds; ++q) A.queue_begin[q] = b->queue_begin[q];
    A.work = b->d_work;
    A.frame = b->d_frame;
    A.spill = b->d_spill;
    A.spill_depth = b->spill_depth;
    A.threads = b->cus * p.blocks_per_cu * kBlock;
    A.rec = b->d_rec;
    A.stats = stats;
    A.tile_w = p.tile_w;
    A.tile_shift = log2u(p.tile_w);
    A.fetch = p.tiles_per_fetch;
    A.nscenes = (uint32_t)b->scenes.size();
    return A;
}

BatchLaunch batch_launch(const dmt_batch* b, const dmt_batch_params& p) {
    bool feat = false;
    uint32_t ext = 0u;
    for (size_t i = 0; i < b->frames.size(); ++i) {
        feat = feat || uses_features(b->scenes[i], &b->frames[i]);
        ext |= ext_bits(b->scenes[i], &b->frames[i]);
    }
    return BatchLaunch{b->cus * p.blocks_per_cu, p.stack_lds, feat, ext};
}

void drop_graph(dmt_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

dmt_status upload_table(dmt_batch* b, BatchScene* dst, const std::vector<BatchScene>& t) {
    hipError_t e = hipMemcpyAsync(dst, t.data(), t.size() * sizeof(BatchScene), hipMemcpyHostToDevice,
                                  b->dev->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->dev->stream);
    if (e != hipSuccess) return mark_lost_if_fatal(b->dev, e, "batch table upload");
    return DMT_OK;
}

dmt_status apply_params(dmt_batch* b, const dmt_batch_params& p) {
    dmt_status st = upload_table(b, b->d_scenes, batch_table(b, p.tile_w, nullptr));
    if (st) return st;
    drop_graph(b);
    b->params = p;
    return DMT_OK;
}

dmt_status ensure_graph(dmt_batch* b) {
    if (b->exec) return DMT_OK;
    hipStream_t stream = b->dev->stream;
    const BatchArgs A = batch_args(b, b->params, b->d_scenes, b->d_stats);
    const BatchLaunch L = batch_launch(b, b->params);
    hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    const hipError_t le = launch_batch_frame(A, L, stream);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(stream, &g);
    if (le != hipSuccess || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(le != hipSuccess ? le : e, "batch frame capture");
    }
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        b->exec = nullptr;
        (void)hipGraphDestroy(g);
        return hip_fail(e, "hipGraphInstantiate");
    }
    b->graph = g;
    return DMT_OK;
}

void batch_free(dmt_batch* b) {
    (void)hipSetDevice(b->dev->id);
    (void)sync_all(b->dev);
    drop_graph(b);
    void* bufs[] = {b->d_scenes, b->d_order, b->d_work, b->d_frame, b->d_rec, b->d_spill, b->d_stats};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    for (dmt_scene* s : b->scenes) {
        if (--s->batches == 0 && s->destroy_requested) scene_free(s);
    }
    delete b;
}

}  // namespace

extern "C" {

dmt_status dmt_batch_create(dmt_device* dev, dmt_scene* const* scenes, const dmt_frame_params* frames,
                            uint32_t count, const dmt_batch_params* params, dmt_batch** out) {
    if (!out) return fail(DMT_ERR_INVALID_ARG, "out is null");
    *out = nullptr;
    dmt_status st = check_device(dev);
    if (st) return st;
    if (count == 0 || !scenes || !frames) return fail(DMT_ERR_INVALID_ARG, "a batch needs scenes and frames");
    if (count > kBatchMaxScenes) return fail(DMT_ERR_LIMIT, "at most 4096 scenes per batch");
    dmt_batch_params p;
    if ((st = normalize_params(params, &p))) return st;
    std::vector<dmt_scene*> list(scenes, scenes + count);
    for (uint32_t i = 0; i < count; ++i) {
        if (!list[i]) return fail(DMT_ERR_INVALID_ARG, "scene " + std::to_string(i) + " is null");
        if (list[i]->dev != dev) return fail(DMT_ERR_INVALID_ARG, "scene " + std::to_string(i) + " is on another device");
        if ((st = check_frame_for(list[i], &frames[i]))) return fail(st, "frame " + std::to_string(i) + ": " + g_err);
    }
    {
        std::vector<dmt_scene*> sorted = list;
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            return fail(DMT_ERR_INVALID_ARG, "a scene appears twice in the batch");
    }
    for (uint32_t i = 0; i < count; ++i) {
        bool resized = false;
        if ((st = ensure_accum(list[i], frames[i].width, frames[i].height, &resized)))
            return fail(st, "scene " + std::to_string(i) + ": " + g_err);
    }
    int cus = 0;
    hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev->id);
    if (e != hipSuccess || cus <= 0) return hip_fail(e, "hipDeviceGetAttribute(CU count)");

    auto* b = new (std::nothrow) dmt_batch();
    if (!b) return fail(DMT_ERR_OUT_OF_MEMORY, "host allocation");
    b->dev = dev;
    b->scenes = list;
    b->frames.assign(frames, frames + count);
    b->cus = (uint32_t)cus;
    b->max_threads = b->cus * kBatchMaxBlocksPerCu * kBlock;
    for (dmt_scene* s : list) s->batches++;
    uint32_t need = 0;
    for (dmt_scene* s : list) need = std::max(need, 3u * s->bvh_depth + 4u);
    b->spill_depth = need > 16u ? need - 16u : 0u;

    // Scenes to XCD queues: largest first onto the least loaded queue.
    std::vector<uint32_t> idx(count);
    for (uint32_t i = 0; i < count; ++i) idx[i] = i;
    auto cost = [&](uint32_t i) { return (uint64_t)frames[i].width * frames[i].height * frames[i].spp; };
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t c) { return cost(a) > cost(c); });
    std::vector<std::vector<uint32_t>> queue(kXcds);
    uint64_t load[kXcds] = {};
    for (uint32_t i : idx) {
        uint32_t q = 0;
        for (uint32_t k = 1; k < kXcds; ++k)
            if (load[k] < load[q]) q = k;
        queue[q].push_back(i);
        load[q] += cost(i);
    }
    for (uint32_t q = 0; q < kXcds; ++q) {
        b->queue_begin[q] = (uint32_t)b->order.size();
        b->order.insert(b->order.end(), queue[q].begin(), queue[q].end());
    }
    b->queue_begin[kXcds] = (uint32_t)b->order.size();
    // records of the sky lights (sun or HDRI cluster, moon) when a frame has them
    bool sky = false;
    for (uint32_t i = 0; i < count; ++i) sky = sky || sky_lights(list[i], &frames[i]);
    const std::pair<void**, size_t> req[] = {
        {(void**)&b->d_scenes, count * sizeof(BatchScene)},
        {(void**)&b->d_order, count * sizeof(uint32_t)},
        {(void**)&b->d_work, count * sizeof(uint32_t)},
        {(void**)&b->d_frame, sizeof(uint32_t)},
        {(void**)&b->d_rec, (size_t)(sky ? 4u : 2u) * kRecRows * b->max_threads * sizeof(float4)},
        {(void**)&b->d_spill, (size_t)b->max_threads * b->spill_depth * sizeof(uint32_t)},
        {(void**)&b->d_stats, sizeof(DeviceStats)},
    };
    for (const auto& r : req) {
        if (r.second == 0) continue;
        e = hipMalloc(r.first, r.second);
        if (e != hipSuccess) {
            *r.first = nullptr;
            batch_free(b);
            return hip_fail(e, "hipMalloc(batch)");
        }
    }
    e = hipMemcpy(b->d_order, b->order.data(), count * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemsetAsync(b->d_stats, 0, sizeof(DeviceStats), dev->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b->d_frame, 0, sizeof(uint32_t), dev->stream);
    if (e != hipSuccess) {
        batch_free(b);
        return hip_fail(e, "batch set-up");
    }
    if ((st = apply_params(b, p))) {
        batch_free(b);
        return st;
    }
    *out = b;
    return DMT_OK;
}

void dmt_batch_destroy(dmt_batch* b) {
    if (b) batch_free(b);
}

dmt_status dmt_batch_set_params(dmt_batch* b, const dmt_batch_params* params) {
    if (!b || !params) return fail(DMT_ERR_INVALID_ARG, "null pointer");
    dmt_status st = check_device(b->dev);
    if (st) return st;
    dmt_batch_params p;
    if ((st = normalize_params(params, &p))) return st;
    hipError_t e = sync_all(b->dev);
    if (e != hipSuccess) return mark_lost_if_fatal(b->dev, e, "synchronize");
    return apply_params(b, p);
}

dmt_status dmt_batch_get_params(const dmt_batch* b, dmt_batch_params* out) {
    if (!b || !out) return fail(DMT_ERR_INVALID_ARG, "null pointer");
    *out = b->params;
    return DMT_OK;
}

dmt_status dmt_batch_render(dmt_batch* b, uint32_t first_frame, uint32_t frame_count) {
    if (!b) return fail(DMT_ERR_INVALID_ARG, "batch is null");
    dmt_status st = check_device(b->dev);
    if (st) return st;
    if (frame_count == 0) return fail(DMT_ERR_INVALID_ARG, "frame_count must be >= 1");
    hipStream_t stream = b->dev->stream;
    hipError_t e = hipSuccess;
    // after the members' per-scene frames still in flight
    for (dmt_scene* s : b->scenes)
        if (s->last_set > 0 && e == hipSuccess) e = hipStreamWaitEvent(stream, s->acc_done[s->last_set], 0);
    if (e != hipSuccess) return mark_lost_if_fatal(b->dev, e, "hipStreamWaitEvent");
    if (!b->params.plain_launches && (st = ensure_graph(b))) return st;
    e = batch_set_frame(b->d_frame, first_frame, stream);
    const BatchArgs A = batch_args(b, b->params, b->d_scenes, b->d_stats);
    const BatchLaunch L = batch_launch(b, b->params);
    for (uint32_t k = 0; k < frame_count && e == hipSuccess; ++k)
        e = b->params.plain_launches ? launch_batch_frame(A, L, stream) : hipGraphLaunch(b->exec, stream);
    // later per-scene frames wait for the batch's accumulation
    for (dmt_scene* s : b->scenes) {
        if (e == hipSuccess) e = hipEventRecord(s->acc_done[0], stream);
        s->last_set = 0;
    }
    if (e != hipSuccess) return mark_lost_if_fatal(b->dev, e, "batch render");
    b->frames_rendered += frame_count;
    return DMT_OK;
}

dmt_status dmt_batch_autotune(dmt_batch* b, uint32_t frames_per_trial, dmt_batch_params* out_best,
                              float* out_ms) {
    if (!b) return fail(DMT_ERR_INVALID_ARG, "batch is null");
    dmt_status st = check_device(b->dev);
    if (st) return st;
    if (frames_per_trial == 0) frames_per_trial = 2;
    hipStream_t stream = b->dev->stream;
    hipError_t e = sync_all(b->dev);
    if (e != hipSuccess) return mark_lost_if_fatal(b->dev, e, "synchronize");
    // scratch accumulation buffers, table and counters
    size_t pixels = 0;
    for (const auto& f : b->frames) pixels += (size_t)f.width * f.height;
    float4* acc = nullptr;
    BatchScene* table = nullptr;
    DeviceStats* stats = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    e = hipMalloc((void**)&acc, pixels * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc((void**)&table, b->scenes.size() * sizeof(BatchScene));
    if (e == hipSuccess) e = hipMalloc((void**)&stats, sizeof(DeviceStats));
    if (e == hipSuccess) e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    auto release = [&]() {
        if (acc) (void)hipFree(acc);
        if (table) (void)hipFree(table);
        if (stats) (void)hipFree(stats);
        for (hipEvent_t x : ev)
            if (x) (void)hipEventDestroy(x);
    };
    if (e != hipSuccess) {
        release();
        return hip_fail(e, "autotune scratch");
    }
    std::vector<float4*> acc_of(b->scenes.size());
    {
        size_t off = 0;
        for (size_t i = 0; i < b->frames.size(); ++i) {
            acc_of[i] = acc + off;
            off += (size_t)b->frames[i].width * b->frames[i].height;
        }
    }
    auto trial = [&](const dmt_batch_params& p, float* ms) -> dmt_status {
        dmt_status ts = upload_table(b, table, batch_table(b, p.tile_w, &acc_of));
        if (ts) return ts;
        const BatchArgs A = batch_args(b, p, table, stats);
        const BatchLaunch L = batch_launch(b, p);
        hipError_t te = batch_set_frame(b->d_frame, 0u, stream);
        if (te == hipSuccess) te = launch_batch_frame(A, L, stream);   // warm-up
        if (te == hipSuccess) te = hipEventRecord(ev[0], stream);
        for (uint32_t k = 0; k < frames_per_trial && te == hipSuccess; ++k) te = launch_batch_frame(A, L, stream);
        if (te == hipSuccess) te = hipEventRecord(ev[1], stream);
        if (te == hipSuccess) te = hipEventSynchronize(ev[1]);
        float t = 0.0f;
        if (te == hipSuccess) te = hipEventElapsedTime(&t, ev[0], ev[1]);
        if (te != hipSuccess) return mark_lost_if_fatal(b->dev, te, "autotune trial");
        *ms = t / (float)frames_per_trial;
        return DMT_OK;
    };
    dmt_batch_params best = b->params;
    float best_ms = 0.0f;
    st = trial(best, &best_ms);
    // coordinate descent, one parameter at a time
    const uint32_t blocks_v[] = {1, 2, 3, 4, 5}, stack_v[] = {16, 24, 32}, tile_v[] = {8, 16, 32, 64},
                   fetch_v[] = {1, 2, 4, 8};
    struct Axis {
        uint32_t dmt_batch_params::*field;
        const uint32_t* values;
        size_t n;
    };
    const Axis axes[] = {{&dmt_batch_params::blocks_per_cu, blocks_v, 5},
                         {&dmt_batch_params::stack_lds, stack_v, 3},
                         {&dmt_batch_params::tile_w, tile_v, 4},
                         {&dmt_batch_params::tiles_per_fetch, fetch_v, 4}};
    for (const Axis& ax : axes) {
        if (st) break;
        const dmt_batch_params base = best;
        for (size_t k = 0; k < ax.n && !st; ++k) {
            if (ax.values[k] == base.*(ax.field)) continue;
            dmt_batch_params q = base;
            q.*(ax.field) = ax.values[k];
            float ms = 0.0f;
            st = trial(q, &ms);
            if (!st && ms < best_ms) {
                best_ms = ms;
                best = q;
            }
        }
    }
    release();
    if (st) return st;
    if ((st = apply_params(b, best))) return st;
    if (out_best) *out_best = best;
    if (out_ms) *out_ms = best_ms;
    return DMT_OK;
}

dmt_status dmt_batch_stats(dmt_batch* b, dmt_stats* out) {
    if (!b || !out) return fail(DMT_ERR_INVALID_ARG, "null pointer");
    dmt_status st = check_device(b->dev);
    if (st) return st;
    std::memset(out, 0, sizeof(*out));
    DeviceStats h{};
    hipError_t e = sync_all(b->dev);
    if (e == hipSuccess) e = hipMemcpy(&h, b->d_stats, sizeof(h), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return mark_lost_if_fatal(b->dev, e, "batch stats");
    out->primary_rays = h.primary;
    out->extension_rays = h.extension;
    out->shadow_rays = h.shadow;
    out->frames = b->frames_rendered;
    return DMT_OK;
}

dmt_status dmt_batch_reset_stats(dmt_batch* b) {
    if (!b) return fail(DMT_ERR_INVALID_ARG, "batch is null");
    dmt_status st = check_device(b->dev);
    if (st) return st;
    hipError_t e = hipMemsetAsync(b->d_stats, 0, sizeof(DeviceStats), b->dev->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->dev->stream);
    if (e != hipSuccess) return mark_lost_if_fatal(b->dev, e, "batch stats reset");
    b->frames_rendered = 0;
    return DMT_OK;
}

}  // extern "C"
