// C ABI of the gfx950 path tracer (declared in include/dmt.h).
//
// The device owns one HIP stream; a scene owns its HBM buffers (acceleration
// data, wavefront queues, accumulation buffer).  Every call validates its
// arguments and reports failures through a status code and dmt_last_error(),
// as the reference's software device reports through a false/0 return and
// LOG_ERROR (src/rhi_software/SoftwareDevice.cpp:337-421, 539-558).
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/dmt.h"
#include "host/prep.h"
#include "kernels/wavefront.h"

using namespace dmt;

// Frames of per-frame dmt_render calls kept in flight (one stream each).
constexpr uint32_t kFrameSets = 3;

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

struct dmt_device {
    int id = 0;
    hipStream_t stream = nullptr;   // exported stream; frame set 0, uploads, queries
    hipStream_t aux[kFrameSets - 1] = {};   // frame sets 1.. (frames in flight)
    bool lost = false;
    // Scenes hold the device: dmt_device_destroy with live scenes defers the
    // teardown to the last dmt_scene_destroy.
    uint32_t scenes = 0;
    bool destroy_requested = false;
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
    // k % kFrameSets and its stream, so the short late-bounce launches of one
    // frame overlap the next frames' heavy bounces.  Only the final
    // accumulation is ordered frame after frame.
    Wavefront wf[kFrameSets]{};
    std::vector<void*> wf_buffers;
    float4* accum = nullptr;
    uint32_t acc_w = 0, acc_h = 0;
    DeviceStats* stats = nullptr;
    DeviceStats* count_stats = nullptr;
    hipEvent_t acc_done[kFrameSets] = {};
    hipEvent_t poll[kFrameSets][kPollSlots] = {};
    uint32_t* host_count = nullptr;   // pinned, kFrameSets * kPollSlots
    int next_set = 0;
    int last_set = -1;
    // profiling
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    double extend_ms = 0.0;
    uint64_t extend_launches = 0;
    uint64_t frames = 0;
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

template <class T>
dmt_status upload(dmt_scene* s, const T* src, size_t count, const T** out) {
    *out = nullptr;
    if (count == 0) return DMT_OK;
    T* d = nullptr;
    hipError_t e = dalloc(&d, count);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(scene)");
    s->buffers.push_back(d);
    s->scene_bytes += count * sizeof(T);
    e = hipMemcpy(d, src, count * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(scene)");
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

void free_wavefront(dmt_scene* s) {
    for (void* p : s->wf_buffers) (void)hipFree(p);
    s->wf_buffers.clear();
    for (auto& w : s->wf) w = Wavefront{};
    s->accum = nullptr;
    s->acc_w = s->acc_h = 0;
    s->last_set = -1;
    s->next_set = 0;
}

dmt_status ensure_wavefront(dmt_scene* s, uint32_t w, uint32_t h, bool* resized) {
    *resized = false;
    if (s->accum && s->acc_w == w && s->acc_h == h) return DMT_OK;
    hipError_t se = sync_all(s->dev);
    if (se != hipSuccess) return mark_lost_if_fatal(s->dev, se, "synchronize before resize");
    free_wavefront(s);
    *resized = true;
    const size_t n = (size_t)w * h;
    std::vector<std::pair<void**, size_t>> req = {{(void**)&s->accum, n * 16}};
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
            {(void**)&W.counters, kCounterSlots * 4}, {(void**)&W.ps.meta, n * 4},
            {(void**)&W.finished, n * 4}, {(void**)&W.regen, n},
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
            free_wavefront(s);
            return hip_fail(e, "hipMalloc(wavefront)");
        }
        s->wf_buffers.push_back(*r.first);
    }
    hipError_t e = hipMemsetAsync(s->accum, 0, n * 16, s->dev->stream);
    for (int set = 0; set < (int)kFrameSets && e == hipSuccess; ++set)
        e = hipMemsetAsync(s->wf[set].regen, 0, n, s->dev->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->dev->stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemsetAsync");
    s->acc_w = w;
    s->acc_h = h;
    return DMT_OK;
}

dmt_status check_frame(const dmt_frame_params* f) {
    if (!f) return fail(DMT_ERR_INVALID_ARG, "frame params are null");
    if (f->width == 0 || f->height == 0) return fail(DMT_ERR_INVALID_ARG, "width and height must be > 0");
    if ((uint64_t)f->width * f->height > (1ull << 26)) return fail(DMT_ERR_LIMIT, "more than 2^26 pixels");
    if (f->spp < 1 || f->spp > DMT_MAX_SPP) return fail(DMT_ERR_INVALID_ARG, "spp must be in 1..32");
    if (f->max_bounces > DMT_MAX_BOUNCES) return fail(DMT_ERR_LIMIT, "max_bounces must be <= 64");
    if (f->refract_bounces > 64) return fail(DMT_ERR_LIMIT, "refract_bounces must be <= 64");
    return DMT_OK;
}

FrameArgs frame_args(const dmt_frame_params* f) {
    return make_frame_args(f->pos_fovtan, f->fwd_aspect, f->right_xyz, f->up_xyz, f->width, f->height,
                           f->frame_index, f->spp, f->max_bounces, f->reset_accum, f->mis_enabled,
                           f->light_tree_enabled, f->refract_bounces, f->firefly_clamp, f->ema_alpha);
}

dmt_status harvest_events(dmt_scene* s) {
    if (s->ev_used == 0) return DMT_OK;
    hipError_t e = sync_all(s->dev);
    if (e != hipSuccess) return mark_lost_if_fatal(s->dev, e, "synchronize");
    for (size_t k = 0; k + 1 < s->ev_used; k += 2) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, s->ev_pool[k], s->ev_pool[k + 1]) == hipSuccess) s->extend_ms += ms;
        s->extend_launches++;
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
    e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
    for (uint32_t k = 0; k + 1 < kFrameSets && e == hipSuccess; ++k)
        e = hipStreamCreateWithFlags(&d->aux[k], hipStreamNonBlocking);
    if (e != hipSuccess) {
        device_free(d);
        return hip_fail(e, "hipStreamCreate");
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
    PreparedScene P;
    std::string msg;
    try {
        msg = prepare_scene(desc->prims, desc->prim_count, desc->lights, desc->light_count, desc->positions,
                            desc->vertex_count, desc->indices, desc->index_count, desc->env_rgb, desc->env_width,
                            desc->env_height, desc->env_intensity, bp ? bp->sah_bins : 0u,
                            bp ? bp->max_leaf : 0u, P);
    } catch (const std::bad_alloc&) {
        return fail(DMT_ERR_OUT_OF_MEMORY, "host allocation during scene build");
    }
    if (!msg.empty()) return fail(DMT_ERR_INVALID_ARG, msg);
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
    V.prim_count = P.prim_count;
    V.plane_count = (uint32_t)P.plane_ids.size();
    if ((st = upload(s, (const float4*)P.planes.data(), P.plane_ids.size(), &V.planes))) return bail(st);
    if ((st = upload(s, P.plane_ids.data(), P.plane_ids.size(), &V.plane_ids))) return bail(st);
    if ((st = upload(s, (const float4*)P.sphere_bvh.nodes.data(), P.sphere_bvh.nodes.size() * 8u, &f4))) return bail(st);
    V.sph_nodes = f4;
    if ((st = upload(s, (const float4*)P.spheres.data(), P.sphere_ids.size(), &V.spheres))) return bail(st);
    if ((st = upload(s, P.sphere_ids.data(), P.sphere_ids.size(), &V.sphere_ids))) return bail(st);
    if ((st = upload(s, (const float4*)P.tri_bvh.nodes.data(), P.tri_bvh.nodes.size() * 8u, &f4))) return bail(st);
    V.tri_nodes = f4;
    if ((st = upload(s, (const float4*)P.tris.data(), (size_t)P.tri_count * 3u, &V.tris))) return bail(st);
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
    s->bvh_depth = std::max(P.tri_bvh.depth, P.sphere_bvh.depth);
    s->bvh_nodes = (uint32_t)(P.tri_bvh.nodes.size() + P.sphere_bvh.nodes.size());
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

void dmt_scene_destroy(dmt_scene* s) {
    if (!s) return;
    if (s->dev) {
        (void)hipSetDevice(s->dev->id);
        (void)sync_all(s->dev);
    }
    free_wavefront(s);
    for (hipEvent_t e : s->acc_done)
        if (e) (void)hipEventDestroy(e);
    for (auto& row : s->poll)
        for (hipEvent_t e : row)
            if (e) (void)hipEventDestroy(e);
    if (s->host_count) (void)hipHostFree(s->host_count);
    for (void* p : s->buffers) (void)hipFree(p);
    for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
    dmt_device* d = s->dev;
    delete s;
    if (d && --d->scenes == 0 && d->destroy_requested) device_free(d);
}

namespace {

// Renders frame_count consecutive frames, each on the next wavefront set and
// its stream: frames stay in flight while the accumulation passes are ordered
// frame after frame by events.
dmt_status render_impl(dmt_scene* s, const dmt_frame_params* f, uint32_t frame_count) {
    if (!s) return fail(DMT_ERR_INVALID_ARG, "scene is null");
    dmt_status st = check_device(s->dev);
    if (st) return st;
    if ((st = check_frame(f))) return st;
    if (frame_count == 0) return fail(DMT_ERR_INVALID_ARG, "frame_count must be >= 1");
    bool resized = false;
    if ((st = ensure_wavefront(s, f->width, f->height, &resized))) return st;
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
            o.extend_events = s->ev_pool.data() + s->ev_used;
            o.max_events = (uint32_t)bound;
        }
        hipError_t e = launch_frame(s->view, F, s->wf[set], s->accum, stream, o);
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

dmt_status dmt_trace_closest(dmt_scene* s, const float* rays, uint32_t n, dmt_hit* out) {
    static_assert(sizeof(dmt_hit) == sizeof(HitOut), "dmt_hit layout");
    if (!s || (n && (!rays || !out))) return fail(DMT_ERR_INVALID_ARG, "null pointer");
    dmt_status st = check_device(s->dev);
    if (st) return st;
    if (n == 0) return DMT_OK;
    bool resized = false;
    if (!s->accum && (st = ensure_wavefront(s, 1, 1, &resized))) return st;
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
    s->frames = 0;
    {
        hipError_t e = hipMemsetAsync(s->stats, 0, sizeof(DeviceStats), s->dev->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->dev->stream);
        if (e != hipSuccess) return hip_fail(e, "hipMemsetAsync");
    }
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
    if ((st = check_frame(f))) return st;
    bool resized = false;
    if ((st = ensure_wavefront(s, f->width, f->height, &resized))) return st;
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
    if (e == hipSuccess) e = launch_frame(s->view, F, W, scratch, s->dev->stream, o);
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
    out->frames = 1;
    out->scene_bytes = s->scene_bytes;
    out->bvh_nodes = s->bvh_nodes;
    out->bvh_depth = s->bvh_depth;
    return DMT_OK;
}

}  // extern "C"
