// HipDevice: the HIP path tracer as a render-hardware-interface device.
// Declarations and the behaviour of each call: include/dmt_rhi.h.
//
// Resources follow the reference's software device (SoftwareDevice.cpp):
// buffers and acceleration structures are host copies (the scene build reads
// them on the host, as Embree reads the software device's copies); textures
// are device memory, RGBA32F for every format (the software device's float
// backing, SoftwareDevice.cpp:485-537), with the last host upload kept for
// the env map.  A pathtrace dispatch renders with the C ABI of dmt.h.
#include "../../../include/dmt_rhi.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <new>
#include <unordered_map>
#include <vector>

namespace dmt::rhi {

namespace {

constexpr std::uint32_t kPbrNoTexTile = 0xFFFFFFFFu;   // PathTrace.slang:317
constexpr std::size_t kPrimFloats = 28;                 // 7 float4 per primitive (Engine.cpp:4924-4994)
constexpr std::size_t kLightFloats = 20;                // 5 float4 per light (Engine.cpp:5235-5300)

float f32(const std::uint8_t* p, std::size_t off) {
    float v;
    std::memcpy(&v, p + off, 4);
    return v;
}

std::uint32_t u32(const std::uint8_t* p, std::size_t off) {
    std::uint32_t v;
    std::memcpy(&v, p + off, 4);
    return v;
}

std::uint32_t bits(float v) {
    std::uint32_t u;
    std::memcpy(&u, &v, 4);
    return u;
}

}  // namespace

bool DecodePathtracePush(const void* data, std::size_t size, bool strict, PathtracePush* out, std::string* error) {
    auto fail = [error](std::string m) {
        if (error) *error = std::move(m);
        return false;
    };
    if (!out) return fail("out is null");
    if (size && !data) return fail("push data is null");
    if (size < ptpush::kHeadBytes)
        return fail("push of " + std::to_string(size) + " B is shorter than the 112 B PtPush head");
    std::array<std::uint8_t, slot::kMaxPushBytes> b{};   // fields past the pushed size read as zero
    std::memcpy(b.data(), data, std::min(size, b.size()));
    const std::uint8_t* p = b.data();
    using namespace ptpush;

    PathtracePush r;
    dmt_frame_params& F = r.frame;
    for (int i = 0; i < 4; ++i) {
        F.pos_fovtan[i] = f32(p, kPosFovTan + 4 * i);
        F.fwd_aspect[i] = f32(p, kFwdAspect + 4 * i);
        F.right_xyz[i] = f32(p, kRightXyz + 4 * i);
        F.up_xyz[i] = f32(p, kUpXyz + 4 * i);
    }
    F.width = F.height = 0;
    F.frame_index = u32(p, kFrameIndex);
    F.reset_accum = u32(p, kResetAccum) != 0u;
    F.max_bounces = u32(p, kMaxBounces);
    F.spp = u32(p, kSpp);
    F.mis_enabled = u32(p, kMisEnabled) != 0u;
    F.light_tree_enabled = u32(p, kLightTreeEnabled) != 0u;
    // int(w2j_row1.w), the shadow ray's dielectric budget (PathTrace.slang:5100;
    // the engine writes 0..16, Engine.cpp:8676)
    const float rb = f32(p, kW2jRow1 + 12);
    F.refract_bounces = rb >= 1.0f ? (std::uint32_t)std::min(rb, 64.0f) : 0u;
    F.firefly_clamp = f32(p, kW2jRow2 + 12);   // PathTrace.slang:3366-3370
    F.ema_alpha = f32(p, kAccumParams);        // PathTrace.slang:7765
    F.flags = 0u;
    r.tlas_present = u32(p, kTlasPresent) != 0u;
    r.prim_count = u32(p, kPrimCount);
    r.light_count = u32(p, kLightCount);
    r.env_map_present = u32(p, kEnvMapPresent) != 0u;
    r.env_intensity = f32(p, kEnvIntensity);
    r.env_total_luminance = f32(p, kEnvTotalLuminance);
    r.tonemap_op = u32(p, kTonemapParams);
    // camera-sample features: thin lens (PathTrace.slang:5786-5816), shutter
    // time (:5771), hero wavelength (:5756), mesh translation (:2483)
    for (int i = 0; i < 4; ++i) F.dof[i] = f32(p, kDofParams + 4 * i);
    F.cat_eye = f32(p, kDofBokehParams);
    r.motion_blur = f32(p, kMotionBlurParams) > 0.5f;
    F.motion_blur = r.motion_blur;
    F.spectral = f32(p, kSpectralParams) > 0.5f;
    F.cauchy_c = f32(p, kSpectralParams + 4);
    for (int i = 0; i < 3; ++i) {
        F.mesh_shift_prev[i] = f32(p, kMeshMotionPrev + 4 * i);
        F.mesh_shift_curr[i] = f32(p, kMeshMotionCurr + 4 * i);
    }
    F.mesh_shift_active = f32(p, kMeshMotionPrev + 12) > 0.5f;
    r.denoiser = u32(p, kDenoiserEnabled) != 0u;
    // with a denoiser every pixel's first sample takes the frame's Halton
    // offset (PathTrace.slang:5730-5731)
    if (r.denoiser) {
        F.flags |= DMT_FRAME_FIXED_JITTER;
        F.jitter[0] = f32(p, kHaltonJitter);
        F.jitter[1] = f32(p, kHaltonJitter + 4);
    }
    // sky lights: the gradient sky's sun and moon, the env map's cluster lights
    for (int i = 0; i < 4; ++i) {
        F.sun[i] = i < 3 ? f32(p, kSunAndMode + 4 * i) : 0.0f;
        F.sun_extra[i] = i < 3 ? f32(p, kSunExtra + 4 * i) : 0.0f;
        F.moon[i] = f32(p, kMoonDirPhase + 4 * i);
        F.moon_extra[i] = i < 3 ? f32(p, kMoonExtra + 4 * i) : 0.0f;
    }
    F.sun_disc_brightness = f32(p, kSunExtra2);
    const std::uint32_t hdri = u32(p, kHdriLightsCount);
    F.hdri_light_count = r.env_map_present ? std::min<std::uint32_t>(hdri, 8u) : 0u;
    for (std::uint32_t i = 0; i < 8u; ++i)
        for (int c = 0; c < 4; ++c) {
            F.hdri_dir[i][c] = i < F.hdri_light_count ? f32(p, kHdriLightsDir + 16 * i + 4 * c) : 0.0f;
            F.hdri_col[i][c] = i < F.hdri_light_count ? f32(p, kHdriLightsCol + 16 * i + 4 * c) : 0.0f;
        }
    r.write_normal = u32(p, kWriteNormalGbuffer) != 0u;
    r.write_albedo = u32(p, kWriteAlbedoGbuffer) != 0u;
    r.write_specular_albedo = u32(p, kWriteSpecularAlbedo) != 0u;
    r.write_roughness = u32(p, kWriteRoughness) != 0u;
    r.write_specular_distance = u32(p, kWriteSpecularHitDistance) != 0u;
    r.write_hdr_aux = u32(p, kWriteHdrAux) != 0u;
    r.hdr_pipeline = f32(p, kW2jRow0 + 12) > 0.5f;   // r_hdr_pipeline (Engine.cpp:8667-8669)
    r.clouds_mode = u32(p, kCloudsRuntime);
    for (int i = 0; i < 16; ++i) {
        r.curr_view_proj[i] = f32(p, kCurrViewProj + 4 * i);
        r.prev_view_proj[i] = f32(p, kPrevViewProj + 4 * i);
    }

    std::string feats;
    auto need = [&feats](bool on, const char* what) {
        if (!on) return;
        if (!feats.empty()) feats += ", ";
        feats += what;
    };
    // uint(sun_and_mode.w) selects the sky (PathTrace.slang:3373-3393)
    const float mode_f = f32(p, kSunAndMode + 12);
    const std::uint32_t mode = (mode_f >= 1.0f && mode_f < 4.0f) ? (std::uint32_t)mode_f : 0u;
    need(mode == 2u, "procedural sky (sun_and_mode.w = 2)");
    need(mode == 3u, "Hosek-Wilkie sky (sun_and_mode.w = 3)");
    need(r.env_map_present && hdri > 8u, "more than 8 HDRI cluster lights");
    need(f32(p, kVolParams) > 0.0f || f32(p, kCloudsP1 + 12) > 0.0f || f32(p, kCloudsP3 + 12) > 0.0f,
         "participating media (fog, clouds, Rayleigh)");                                     // :7011
    // SDF clusters (PathTrace.slang:2224-2238): count, step cap, surface distance
    r.sdf_count = u32(p, kSdfParams);
    F.sdf_max_iters = u32(p, kSdfParams + 4);
    F.sdf_epsilon = f32(p, kSdfParamsF);
    need(r.sdf_count != 0u && u32(p, kSdfParams + 8) != 0u, "SDF iteration heat map (sdf_params.z)");
    need(u32(p, kRestirEnabled) != 0u, "ReSTIR");                                            // :6315
    need(f32(p, kOceanParams0) != 0.0f, "FFT ocean");                                         // :6791, :5484
    // water (Engine.cpp:8883-8890, 9263-9266): absorption rgb and index,
    // wave scale / amplitude / speed / time, the camera's medium and the
    // deep-water tint
    for (int k = 0; k < 3; ++k) F.water_absorb[k] = f32(p, kWaterParams0 + 4 * k);
    F.water_ior = f32(p, kWaterParams0 + 12);
    for (int k = 0; k < 4; ++k) F.wave[k] = f32(p, kWaterParams1 + 4 * k);
    F.camera_in_water = f32(p, kWaterState) > 0.5f ? 1u : 0u;
    for (int k = 0; k < 3; ++k) F.deep_color[k] = f32(p, kWaterState + 4 + 4 * k);
    need(u32(p, kMatOverrideCount) != 0u, "material overrides");                             // :1164
    need(f32(p, kSmokeParams) != 0.0f, "smoke volumes");                                     // :981
    need(u32(p, kWriteShadowVis) != 0u, "SIGMA shadow visibility (write_shadow_vis)");       // :5693-5702
    // the denoiser's colour input is the frame's radiance (:7800); this
    // backend forms it from the mean, which is that radiance when the frame
    // resets the accumulation (the engine resets every frame while a denoiser
    // runs, Engine.cpp:7834)
    need(r.denoiser && r.write_hdr_aux && !F.reset_accum, "denoiser colour input without reset_accum");
    // the mesh's PBR map tiles (the engine sets them only with a UV buffer at
    // slot 17, Engine.cpp:9131-9134)
    bool mesh_tex = false;
    for (int k = 0; k < 4; ++k) {
        F.mesh_tex[k] = u32(p, kMeshTexIndices + 4 * k);
        mesh_tex |= F.mesh_tex[k] != kPbrNoTexTile;
    }
    if (mesh_tex) F.flags |= DMT_FRAME_MESH_TEXTURES;
    if (!feats.empty()) {
        if (strict)
            return fail("pathtrace push asks for what this backend does not render: " + feats +
                        " (a lenient device renders without them)");
        r.ignored = feats;
    }
    *out = std::move(r);
    return true;
}

namespace {

struct Buffer {
    std::vector<std::uint8_t> data;
    std::uint64_t gen = 0;
};

struct Texture {
    std::uint32_t w = 0, h = 0;
    TextureFormat format = TextureFormat::Unknown;
    float* dev = nullptr;        // w * h float4
    std::vector<float> host;     // last host upload, while host_valid
    bool host_valid = false;
    std::vector<std::uint8_t> rgba8;   // RGBA8_UNORM: the bytes of the last upload (the PBR atlas)
    std::uint64_t gen = 0;
    std::size_t bytes() const { return (std::size_t)w * h * 16u; }
};

struct Accel {
    bool tlas = false;
    std::vector<float> positions;        // BLAS: its mesh; TLAS: its instances flattened
    std::vector<std::uint32_t> indices;
    std::vector<std::uint32_t> instance_vertices;   // TLAS: the vertex count of each flattened instance
};

// What the current scene was built from.
struct SceneKey {
    std::uint64_t prims = 0, prims_gen = 0, prims_off = 0;
    std::uint32_t prim_count = 0;
    std::uint64_t lights = 0, lights_gen = 0, lights_off = 0;
    std::uint32_t light_count = 0;
    std::uint64_t tlas = 0;
    std::uint64_t env = 0, env_gen = 0;
    std::uint32_t env_intensity = 0;
    // the engine's importance CDF (buffer slots 4, 5), read when the env map
    // has extracted cluster lights (their pixels are out of it)
    std::uint64_t cdf_m = 0, cdf_m_gen = 0, cdf_c = 0, cdf_c_gen = 0;
    std::uint32_t cdf_total = 0;
    // the SDF clusters (buffer slot 10)
    std::uint64_t sdf = 0, sdf_gen = 0, sdf_off = 0;
    std::uint32_t sdf_count = 0;
    // the mesh UVs (buffer slot 17), read while the frame gives the mesh map tiles
    std::uint64_t uvs = 0, uvs_gen = 0, uvs_off = 0;
    bool operator==(const SceneKey&) const = default;
};

// Primitive records equal outside the material fields dmt_scene_set_materials
// replaces (albedo rgb + roughness = v1, emission rgb = v4.xyz).
bool same_geometry(const std::vector<float>& a, const std::vector<float>& b) {
    if (a.size() != b.size()) return false;
    for (std::size_t i = 0; i < a.size(); i += kPrimFloats) {
        if (std::memcmp(&a[i], &b[i], 4 * 4) != 0) return false;                   // v0
        if (std::memcmp(&a[i + 8], &b[i + 8], 8 * 4) != 0) return false;           // v2, v3
        if (std::memcmp(&a[i + 19], &b[i + 19], 9 * 4) != 0) return false;         // v4.w, v5, v6
    }
    return true;
}

// Appends BLAS b to the flattened mesh t, each vertex through the row-major 3x4
// transform m; an identity transform keeps the vertices bit for bit.
void append_instance(Accel& t, const Accel& b, const float* m) {
    static const float kIdentity[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    const bool identity = std::memcmp(m, kIdentity, sizeof(kIdentity)) == 0;
    const std::size_t base = t.positions.size() / 3u;
    t.positions.reserve(t.positions.size() + b.positions.size());
    for (std::size_t v = 0; v < b.positions.size(); v += 3) {
        const float x = b.positions[v], y = b.positions[v + 1], z = b.positions[v + 2];
        if (identity) {
            t.positions.insert(t.positions.end(), {x, y, z});
        } else {
            t.positions.insert(t.positions.end(), {m[0] * x + m[1] * y + m[2] * z + m[3],
                                                   m[4] * x + m[5] * y + m[6] * z + m[7],
                                                   m[8] * x + m[9] * y + m[10] * z + m[11]});
        }
    }
    t.indices.reserve(t.indices.size() + b.indices.size());
    for (std::uint32_t ix : b.indices) t.indices.push_back((std::uint32_t)base + ix);
    t.instance_vertices.push_back((std::uint32_t)(b.positions.size() / 3u));
}

}  // namespace

struct HipDevice::Impl {
    HipDeviceOptions opt;
    dmt_device* dev = nullptr;
    hipStream_t stream = nullptr;
    bool lost = false;
    std::string error;
    std::string ignored;
    mutable std::recursive_mutex mu;
    std::uint64_t next_id = 1;
    std::unordered_map<std::uint64_t, Buffer> buffers;
    std::unordered_map<std::uint64_t, Texture> textures;
    std::unordered_map<std::uint64_t, Accel> accels;
    std::unordered_map<std::uint64_t, std::string> pipelines;
    std::unique_ptr<HipCommandBuffer> cmd;
    std::uint64_t output_tex = 0;
    std::uint32_t width = 0, height = 0, frame_index = 0;
    dmt_scene* scene = nullptr;
    SceneKey key{};
    std::vector<float> scene_prims;
    float* scene_accum = nullptr;        // texture memory the scene renders into
    std::uint64_t builds = 0, material_updates = 0, mesh_refits = 0;
    std::vector<std::uint32_t> scene_indices;   // the mesh triangles the scene was built with
    // the moon texture last handed to the scene (dmt_scene_set_moon_map)
    const dmt_scene* moon_scene = nullptr;
    std::uint64_t moon_id = 0, moon_gen = 0;
    // the PBR atlas last handed to the scene (dmt_scene_set_pbr_atlas)
    const dmt_scene* atlas_scene = nullptr;
    std::uint64_t atlas_id = 0, atlas_gen = 0;

    void report(const std::string& m) {
        error = m;
        std::fprintf(stderr, "[dmt_rhi] %s\n", m.c_str());
    }
    // A failed C ABI call: its message, and the device-lost latch.
    bool check(dmt_status st, const char* what) {
        if (st == DMT_OK) return true;
        report(std::string(what) + ": " + dmt_last_error());
        if (dev && dmt_device_is_lost(dev)) lost = true;
        return false;
    }
    bool hip(hipError_t e, const char* what) {
        if (e == hipSuccess) return true;
        report(std::string(what) + ": " + hipGetErrorString(e));
        return false;
    }
    // Orders host access to device memory after every frame in flight.
    bool idle() { return !dev || check(dmt_device_synchronize(dev), "synchronize"); }

    Buffer* buffer(std::uint64_t id) {
        auto it = buffers.find(id);
        return it == buffers.end() ? nullptr : &it->second;
    }
    Texture* texture(std::uint64_t id) {
        auto it = textures.find(id);
        return it == textures.end() ? nullptr : &it->second;
    }
    Accel* accel(std::uint64_t id) {
        auto it = accels.find(id);
        return it == accels.end() ? nullptr : &it->second;
    }

    bool alloc_texture(Texture& t) {
        t.dev = nullptr;
        if (t.bytes() == 0) return true;
        if (!dev) {
            report("no HIP device");
            return false;
        }
        if (!hip(hipSetDevice(opt.device_index), "hipSetDevice")) return false;
        if (!hip(hipMalloc((void**)&t.dev, t.bytes()), "hipMalloc(texture)")) {
            t.dev = nullptr;
            return false;
        }
        return hip(hipMemsetAsync(t.dev, 0, t.bytes(), stream), "hipMemsetAsync(texture)") &&
               hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
    void free_texture(Texture& t) {
        if (!t.dev) return;
        if (t.dev == scene_accum && scene) {
            (void)check(dmt_scene_bind_accum(scene, nullptr, 0, 0), "unbind accumulation texture");
            scene_accum = nullptr;
        }
        (void)idle();
        (void)hipSetDevice(opt.device_index);
        (void)hipFree(t.dev);
        t.dev = nullptr;
    }
    // Host RGBA32F upload into a texture (sizes checked by the caller).
    bool upload(Texture& t, const float* rgba) {
        if (!t.dev) return true;
        if (!idle() || !hip(hipSetDevice(opt.device_index), "hipSetDevice")) return false;
        return hip(hipMemcpyAsync(t.dev, rgba, t.bytes(), hipMemcpyHostToDevice, stream), "texture upload") &&
               hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
    bool host_copy(Texture& t) {
        if (t.host_valid) return true;
        t.host.assign((std::size_t)t.w * t.h * 4u, 0.0f);
        if (t.dev) {
            if (!idle() || !hip(hipSetDevice(opt.device_index), "hipSetDevice")) return false;
            if (!hip(hipMemcpy(t.host.data(), t.dev, t.bytes(), hipMemcpyDeviceToHost), "texture readback"))
                return false;
        }
        t.host_valid = true;
        return true;
    }

    bool ensure_scene(const HipCommandBuffer& cb, const PathtracePush& P);
};

// ---- device ------------------------------------------------------------------

HipDevice::HipDevice(const HipDeviceOptions& options) : impl_(std::make_unique<Impl>()) {
    Impl& I = *impl_;
    I.opt = options;
    I.width = options.width ? options.width : 1280u;
    I.height = options.height ? options.height : 720u;
    I.cmd.reset(new HipCommandBuffer(this));
    if (!I.check(dmt_device_create(options.device_index, &I.dev), "dmt_device_create")) {
        I.dev = nullptr;
        I.lost = true;
        return;
    }
    void* st = nullptr;
    if (I.check(dmt_device_stream(I.dev, &st), "dmt_device_stream")) I.stream = static_cast<hipStream_t>(st);
}

HipDevice::~HipDevice() {
    Impl& I = *impl_;
    if (I.scene) dmt_scene_destroy(I.scene);   // waits for its frames
    I.scene = nullptr;
    I.scene_accum = nullptr;
    for (auto& [id, t] : I.textures) I.free_texture(t);
    if (I.dev) dmt_device_destroy(I.dev);
}

std::unique_ptr<HipDevice> CreateHipDevice(const HipDeviceOptions& options) {
    auto d = std::make_unique<HipDevice>(options);
    if (d->IsDeviceLost()) return nullptr;
    return d;
}

BufferHandle HipDevice::CreateBuffer(const BufferDesc& d) {
    std::lock_guard lock(impl_->mu);
    Buffer b;
    try {
        b.data.assign(d.size, 0);
    } catch (const std::bad_alloc&) {
        impl_->report("CreateBuffer: out of host memory for " + std::to_string(d.size) + " B");
        return BufferHandle{0};
    }
    const std::uint64_t id = impl_->next_id++;
    impl_->buffers.emplace(id, std::move(b));
    return BufferHandle{id};
}

TextureHandle HipDevice::CreateTexture(const TextureDesc& d) {
    std::lock_guard lock(impl_->mu);
    Impl& I = *impl_;
    Texture t;
    t.w = d.width;
    t.h = d.height;
    t.format = d.format;
    if (!I.alloc_texture(t)) return TextureHandle{0};
    const std::uint64_t id = I.next_id++;
    I.textures.emplace(id, std::move(t));
    return TextureHandle{id};
}

PipelineHandle HipDevice::CreateComputePipeline(const ComputePipelineDesc& d) {
    std::lock_guard lock(impl_->mu);
    const std::uint64_t id = impl_->next_id++;
    impl_->pipelines.emplace(id, std::string(d.kernel_name));
    return PipelineHandle{id};
}

AccelStructHandle HipDevice::CreateBLAS(const BLASDesc& d) {
    std::lock_guard lock(impl_->mu);
    Impl& I = *impl_;
    // SoftwareDevice.cpp:337-377: an empty mesh is refused; the arrays are copied
    if (d.vertex_count == 0 || d.index_count == 0 || !d.vertex_positions || !d.indices) {
        I.report("CreateBLAS: empty mesh");
        return AccelStructHandle{0};
    }
    if (d.index_count % 3u) {
        I.report("CreateBLAS: index_count is not a multiple of 3");
        return AccelStructHandle{0};
    }
    for (std::uint32_t i = 0; i < d.index_count; ++i)
        if (d.indices[i] >= d.vertex_count) {
            I.report("CreateBLAS: index " + std::to_string(i) + " is out of range");
            return AccelStructHandle{0};
        }
    Accel a;
    try {
        a.positions.assign(d.vertex_positions, d.vertex_positions + (std::size_t)d.vertex_count * 3u);
        a.indices.assign(d.indices, d.indices + d.index_count);
    } catch (const std::bad_alloc&) {
        I.report("CreateBLAS: out of host memory");
        return AccelStructHandle{0};
    }
    const std::uint64_t id = I.next_id++;
    I.accels.emplace(id, std::move(a));
    return AccelStructHandle{id};
}

AccelStructHandle HipDevice::CreateTLAS(const TLASDesc& d) {
    std::lock_guard lock(impl_->mu);
    Impl& I = *impl_;
    if (d.instances.empty()) return AccelStructHandle{0};
    // The instances flattened into one mesh, each vertex through its row-major
    // 3x4 transform (the software device instances BLAS scenes with the same
    // matrix, SoftwareDevice.cpp:379-421).
    Accel t;
    t.tlas = true;
    try {
        for (const TLASInstance& in : d.instances) {
            const Accel* b = I.accel(in.blas.id);
            if (!b || b->tlas) {
                std::fprintf(stderr, "[dmt_rhi] CreateTLAS: instance references unknown BLAS id=%llu\n",
                             (unsigned long long)in.blas.id);
                continue;
            }
            if (in.mask == 0u) continue;   // never intersected by a ray mask of 0xFF
            if (t.positions.size() / 3u + b->positions.size() / 3u > 0xFFFFFFFFull) {
                I.report("CreateTLAS: more than 2^32 vertices");
                return AccelStructHandle{0};
            }
            append_instance(t, *b, in.transform);
        }
    } catch (const std::bad_alloc&) {
        I.report("CreateTLAS: out of host memory");
        return AccelStructHandle{0};
    }
    const std::uint64_t id = I.next_id++;
    I.accels.emplace(id, std::move(t));
    return AccelStructHandle{id};
}

void HipDevice::DestroyBuffer(BufferHandle h) {
    std::lock_guard lock(impl_->mu);
    impl_->buffers.erase(h.id);
}

void HipDevice::DestroyTexture(TextureHandle h) {
    std::lock_guard lock(impl_->mu);
    Impl& I = *impl_;
    if (Texture* t = I.texture(h.id)) I.free_texture(*t);
    I.textures.erase(h.id);
    if (I.output_tex == h.id) I.output_tex = 0;
}

void HipDevice::DestroyPipeline(PipelineHandle h) {
    std::lock_guard lock(impl_->mu);
    impl_->pipelines.erase(h.id);
}

void HipDevice::DestroyAccelStruct(AccelStructHandle h) {
    std::lock_guard lock(impl_->mu);
    impl_->accels.erase(h.id);
}

void HipDevice::WriteBuffer(BufferHandle h, const void* src, std::size_t size, std::size_t offset) {
    std::lock_guard lock(impl_->mu);
    Impl& I = *impl_;
    Buffer* b = I.buffer(h.id);
    if (!b) return;
    // SoftwareDevice.cpp:539-558: an out-of-range region is refused, not clamped
    if (offset > b->data.size() || size > b->data.size() - offset) {
        I.report("WriteBuffer: copy region (offset " + std::to_string(offset) + " + size " + std::to_string(size) +
                 ") exceeds buffer size " + std::to_string(b->data.size()));
        return;
    }
    if (size) std::memcpy(b->data.data() + offset, src, size);
    b->gen++;
}

bool HipDevice::WriteTexture(TextureHandle h, const void* src, std::size_t size) {
    std::lock_guard lock(impl_->mu);
    Impl& I = *impl_;
    Texture* t = I.texture(h.id);
    if (!t || !src) return false;
    const std::size_t channels = (std::size_t)t->w * t->h * 4u;
    std::vector<float> rgba;
    try {
        rgba.assign(channels, 0.0f);
    } catch (const std::bad_alloc&) {
        I.report("WriteTexture: out of host memory");
        return false;
    }
    if (t->format == TextureFormat::RGBA8_UNORM) {
        // one byte per channel, read as byte / 255 (SoftwareDevice.cpp:497-523)
        if (size != channels) return false;
        const auto* b = static_cast<const std::uint8_t*>(src);
        for (std::size_t i = 0; i < channels; ++i) rgba[i] = (float)b[i] * (1.0f / 255.0f);
        try {
            t->rgba8.assign(b, b + channels);
        } catch (const std::bad_alloc&) {
            I.report("WriteTexture: out of host memory");
            return false;
        }
    } else if (t->format == TextureFormat::RGBA16F) {
        // IEEE binary16 per channel, as the path tracer's shader reads an
        // RGBA16F texture (the engine's moon map, Engine.cpp:4626-4645); the
        // software device, which never samples one, copies the raw bytes
        if (size != channels * 2u) return false;
        const auto* hb = static_cast<const std::uint8_t*>(src);
        for (std::size_t i = 0; i < channels; ++i) {
            const std::uint32_t h16 = (std::uint32_t)hb[2 * i] | ((std::uint32_t)hb[2 * i + 1] << 8);
            const std::uint32_t sign = (h16 & 0x8000u) << 16, e = (h16 >> 10) & 0x1Fu, m = h16 & 0x3FFu;
            float v;
            if (e == 0u) v = std::ldexp((float)m, -24);                                 // zero, subnormal
            else if (e == 31u) v = m ? std::numeric_limits<float>::quiet_NaN() : INFINITY;
            else v = std::ldexp((float)(m | 0x400u), (int)e - 25);
            std::uint32_t vb;
            std::memcpy(&vb, &v, 4);
            vb |= sign;
            std::memcpy(&rgba[i], &vb, 4);
        }
    } else {
        // raw bytes into the float backing, as many as fit (SoftwareDevice.cpp:524-537)
        std::memcpy(rgba.data(), src, std::min(size, channels * 4u));
    }
    if (!I.upload(*t, rgba.data())) return false;
    t->host = std::move(rgba);
    t->host_valid = true;
    t->gen++;
    return true;
}

FrameContext HipDevice::BeginFrame() {
    std::lock_guard lock(impl_->mu);
    Impl& I = *impl_;
    Texture* out = I.texture(I.output_tex);
    if (!out || out->w != I.width || out->h != I.height) {
        if (out) {
            I.free_texture(*out);
            I.textures.erase(I.output_tex);
        }
        Texture t;
        t.w = I.width;
        t.h = I.height;
        t.format = TextureFormat::RGBA32F;
        I.output_tex = 0;
        if (I.alloc_texture(t)) {
            I.output_tex = I.next_id++;
            I.textures.emplace(I.output_tex, std::move(t));
        }
    }
    return FrameContext{TextureHandle{I.output_tex}, I.width, I.height, I.frame_index};
}

void HipDevice::EndFrame(HipCommandBuffer*) {
    std::lock_guard lock(impl_->mu);
    impl_->frame_index++;   // no present: the image stays in the output texture
}

HipCommandBuffer* HipDevice::AcquireCommandBuffer() { return impl_->cmd.get(); }

void HipDevice::Submit(HipCommandBuffer*) {}   // dispatches are already on the device stream

void HipDevice::WaitIdle() {
    std::lock_guard lock(impl_->mu);
    (void)impl_->idle();
}

void HipDevice::Resize(int w, int h) {
    std::lock_guard lock(impl_->mu);
    Impl& I = *impl_;
    I.width = (std::uint32_t)std::max(w, 0);
    I.height = (std::uint32_t)std::max(h, 0);
    // the output texture takes the new size, zeroed (SoftwareDevice.cpp:744-769)
    if (Texture* out = I.texture(I.output_tex)) {
        I.free_texture(*out);
        out->w = I.width;
        out->h = I.height;
        out->host_valid = false;
        out->gen++;
        if (!I.alloc_texture(*out)) out->w = out->h = 0;
    }
}

bool HipDevice::ReadbackTexture(TextureHandle h, void* dst, std::size_t dst_size, std::uint32_t* out_w,
                                std::uint32_t* out_h) {
    std::lock_guard lock(impl_->mu);
    Impl& I = *impl_;
    Texture* t = I.texture(h.id);
    if (!t || !dst) return false;
    if (out_w) *out_w = t->w;
    if (out_h) *out_h = t->h;
    // full or nothing, RGBA32F (SoftwareDevice.cpp:458-483)
    if (dst_size < t->bytes()) return false;
    if (t->host_valid) {
        std::memcpy(dst, t->host.data(), t->bytes());
        return true;
    }
    if (!t->dev) return true;
    if (!I.idle() || !I.hip(hipSetDevice(I.opt.device_index), "hipSetDevice")) return false;
    return I.hip(hipMemcpy(dst, t->dev, t->bytes(), hipMemcpyDeviceToHost), "texture readback");
}

bool HipDevice::ReadbackBuffer(BufferHandle h, void* dst, std::size_t bytes) {
    std::lock_guard lock(impl_->mu);
    Buffer* b = impl_->buffer(h.id);
    if (!b || !dst || bytes > b->data.size()) return false;
    std::memcpy(dst, b->data.data(), bytes);
    return true;
}

const char* HipDevice::DeviceName() const { return "DeMonT HIP path tracer (gfx950)"; }

std::size_t HipDevice::CurrentAllocatedBytes() const {
    std::lock_guard lock(impl_->mu);
    std::size_t n = 0;
    for (const auto& [id, b] : impl_->buffers) n += b.data.size();
    for (const auto& [id, t] : impl_->textures) n += t.bytes() + t.host.size() * 4u;
    for (const auto& [id, a] : impl_->accels) n += a.positions.size() * 4u + a.indices.size() * 4u;
    return n;
}

bool HipDevice::IsDeviceLost() const {
    std::lock_guard lock(impl_->mu);
    return impl_->lost || (impl_->dev && dmt_device_is_lost(impl_->dev));
}

const std::string& HipDevice::LastError() const { return impl_->error; }
dmt_scene* HipDevice::ActiveScene() const { return impl_->scene; }
std::uint64_t HipDevice::SceneBuilds() const { return impl_->builds; }
std::uint64_t HipDevice::MaterialUpdates() const { return impl_->material_updates; }
std::uint64_t HipDevice::MeshRefits() const { return impl_->mesh_refits; }
const std::string& HipDevice::IgnoredFeatures() const { return impl_->ignored; }

// ---- the pathtrace dispatch ------------------------------------------------------

bool HipDevice::Impl::ensure_scene(const HipCommandBuffer& cb, const PathtracePush& P) {
    SceneKey k;
    const Buffer* prims = nullptr;
    const Buffer* lights = nullptr;
    const Accel* tlas = nullptr;
    Texture* env = nullptr;
    const Buffer* cdf_m = nullptr;
    const Buffer* cdf_c = nullptr;
    const Buffer* sdf = nullptr;
    if (P.prim_count) {
        k.prims = cb.buffers_[slot::kPrimBuffer];
        k.prims_off = cb.buffer_offsets_[slot::kPrimBuffer];
        prims = buffer(k.prims);
        if (!prims || k.prims_off > prims->data.size() ||
            (prims->data.size() - k.prims_off) / (kPrimFloats * 4u) < P.prim_count) {
            report("pathtrace: prim_count " + std::to_string(P.prim_count) + " needs " +
                   std::to_string(P.prim_count * kPrimFloats * 4u) + " B at buffer slot 3");
            return false;
        }
        k.prims_gen = prims->gen;
        k.prim_count = P.prim_count;
    }
    if (P.light_count) {
        k.lights = cb.buffers_[slot::kLightBuffer];
        k.lights_off = cb.buffer_offsets_[slot::kLightBuffer];
        lights = buffer(k.lights);
        if (!lights || k.lights_off > lights->data.size() ||
            (lights->data.size() - k.lights_off) / (kLightFloats * 4u) < P.light_count) {
            report("pathtrace: light_count " + std::to_string(P.light_count) + " needs " +
                   std::to_string(P.light_count * kLightFloats * 4u) + " B at buffer slot 12");
            return false;
        }
        k.lights_gen = lights->gen;
        k.light_count = P.light_count;
    }
    const Buffer* uvs = nullptr;
    if (P.tlas_present) {
        k.tlas = cb.accels_[slot::kTlas];
        tlas = accel(k.tlas);
        if (!tlas || !tlas->tlas) {
            report("pathtrace: tlas_present is set but accel slot 2 holds no TLAS");
            return false;
        }
        // the mesh UVs, 2 floats per vertex of the mesh each instance draws
        if ((P.frame.flags & DMT_FRAME_MESH_TEXTURES) && !tlas->indices.empty()) {
            k.uvs = cb.buffers_[slot::kMeshUvBuffer];
            k.uvs_off = cb.buffer_offsets_[slot::kMeshUvBuffer];
            uvs = buffer(k.uvs);
            std::uint32_t need = 0;
            for (std::uint32_t nv : tlas->instance_vertices) need = std::max(need, nv);
            if (!uvs || k.uvs_off > uvs->data.size() || (uvs->data.size() - k.uvs_off) / 8u < need) {
                report("pathtrace: mesh_tex_indices needs the mesh UVs (" + std::to_string(need) +
                       " vertices) at buffer slot 17");
                return false;
            }
            k.uvs_gen = uvs->gen;
        }
    }
    if (P.sdf_count) {
        k.sdf = cb.buffers_[slot::kSdfBuffer];
        k.sdf_off = cb.buffer_offsets_[slot::kSdfBuffer];
        sdf = buffer(k.sdf);
        if (!sdf || k.sdf_off > sdf->data.size() ||
            (sdf->data.size() - k.sdf_off) / (DMT_SDF_CLUSTER_FLOATS * 4u) < P.sdf_count) {
            report("pathtrace: sdf_params.x " + std::to_string(P.sdf_count) + " needs " +
                   std::to_string((std::size_t)P.sdf_count * DMT_SDF_CLUSTER_FLOATS * 4u) + " B at buffer slot 10");
            return false;
        }
        k.sdf_gen = sdf->gen;
        k.sdf_count = P.sdf_count;
    }
    if (P.env_map_present) {
        k.env = cb.textures_[slot::kEnvTexture];
        env = texture(k.env);
        if (!env || env->w == 0 || env->h == 0) {
            report("pathtrace: env_map_present is set but storage-texture slot 5 holds no env map");
            return false;
        }
        k.env_gen = env->gen;
        k.env_intensity = bits(P.env_intensity);
        if (P.frame.hdri_light_count) {
            k.cdf_m = cb.buffers_[slot::kEnvMarginalBuffer];
            k.cdf_c = cb.buffers_[slot::kEnvConditionalBuffer];
            cdf_m = buffer(k.cdf_m);
            cdf_c = buffer(k.cdf_c);
            if (!cdf_m || !cdf_c || cdf_m->data.size() < (std::size_t)env->h * 4u ||
                cdf_c->data.size() < (std::size_t)env->w * env->h * 4u) {
                report("pathtrace: HDRI cluster lights need the env map's CDF at buffer slots 4 (" +
                       std::to_string(env->h) + " floats) and 5 (" + std::to_string((std::size_t)env->w * env->h) +
                       " floats)");
                return false;
            }
            k.cdf_m_gen = cdf_m->gen;
            k.cdf_c_gen = cdf_c->gen;
            k.cdf_total = bits(P.env_total_luminance);
        }
    }
    if (scene && k == key) return true;

    // A new TLAS over the same triangles (instances moved): refit the mesh's
    // tree on the device instead of a rebuild.
    SceneKey only_tlas = key;
    only_tlas.tlas = k.tlas;
    if (scene && tlas && key.tlas && only_tlas == k && !tlas->indices.empty() && tlas->indices == scene_indices) {
        if (!check(dmt_scene_update_mesh(scene, tlas->positions.data(), (std::uint32_t)(tlas->positions.size() / 3u)),
                   "dmt_scene_update_mesh"))
            return false;
        key = k;
        mesh_refits++;
        return true;
    }

    // The primitives as the scene takes them (v3, the previous centre, is the
    // start of the motion-blur shutter; v6 holds the PBR map tiles the atlas
    // at storage-texture slot 16 serves).
    std::vector<float> pr((std::size_t)P.prim_count * kPrimFloats);
    if (P.prim_count) std::memcpy(pr.data(), prims->data.data() + k.prims_off, pr.size() * 4u);

    SceneKey only_prims = key;
    only_prims.prims_gen = k.prims_gen;
    if (scene && P.prim_count && only_prims == k && same_geometry(pr, scene_prims)) {
        std::vector<float> mats((std::size_t)P.prim_count * DMT_MATERIAL_FLOATS);
        for (std::uint32_t i = 0; i < P.prim_count; ++i) {
            const float* q = &pr[(std::size_t)i * kPrimFloats];
            float* m = &mats[(std::size_t)i * DMT_MATERIAL_FLOATS];
            m[0] = q[4]; m[1] = q[5]; m[2] = q[6]; m[3] = q[7];
            m[4] = q[16]; m[5] = q[17]; m[6] = q[18];
        }
        if (!check(dmt_scene_set_materials(scene, mats.data(), mats.size()), "dmt_scene_set_materials")) return false;
        scene_prims = std::move(pr);
        key = k;
        material_updates++;
        return true;
    }

    std::vector<float> env_rgb;
    if (env) {
        if (!host_copy(*env)) return false;
        env_rgb.resize((std::size_t)env->w * env->h * 3u);
        for (std::size_t i = 0; i < (std::size_t)env->w * env->h; ++i)
            for (int c = 0; c < 3; ++c) env_rgb[i * 3 + c] = env->host[i * 4 + c];
    }
    dmt_scene_desc d{};
    d.prims = pr.empty() ? nullptr : pr.data();
    d.prim_count = P.prim_count;
    d.lights = lights ? reinterpret_cast<const float*>(lights->data.data() + k.lights_off) : nullptr;
    d.light_count = P.light_count;
    std::vector<float> uv_flat;
    if (tlas && !tlas->indices.empty()) {
        d.positions = tlas->positions.data();
        d.vertex_count = (std::uint32_t)(tlas->positions.size() / 3u);
        d.indices = tlas->indices.data();
        d.index_count = (std::uint32_t)tlas->indices.size();
        if (uvs) {
            // each flattened instance reads the UVs of its mesh's vertices
            // (the shader indexes mesh_uvs by the mesh's own vertex index)
            uv_flat.reserve((std::size_t)d.vertex_count * 2u);
            const auto* src = reinterpret_cast<const std::uint8_t*>(uvs->data.data() + k.uvs_off);
            for (std::uint32_t nv : tlas->instance_vertices) {
                const std::size_t at = uv_flat.size();
                uv_flat.resize(at + (std::size_t)nv * 2u);
                std::memcpy(uv_flat.data() + at, src, (std::size_t)nv * 8u);
            }
            d.mesh_uvs = uv_flat.data();
        }
    }
    std::vector<float> cdf_copy;
    if (env) {
        d.env_rgb = env_rgb.data();
        d.env_width = env->w;
        d.env_height = env->h;
        d.env_intensity = P.env_intensity;
        if (cdf_m && cdf_c) {   // float-aligned copies of the engine's CDF
            cdf_copy.resize((std::size_t)env->h + (std::size_t)env->w * env->h);
            std::memcpy(cdf_copy.data(), cdf_m->data.data(), (std::size_t)env->h * 4u);
            std::memcpy(cdf_copy.data() + env->h, cdf_c->data.data(), (std::size_t)env->w * env->h * 4u);
            d.env_marginal = cdf_copy.data();
            d.env_conditional = cdf_copy.data() + env->h;
            d.env_total_luminance = P.env_total_luminance;
        }
    }
    // the cluster and light records are copied out of the byte buffers so they are float-aligned
    std::vector<float> sdf_copy;
    if (sdf) {
        sdf_copy.resize((std::size_t)P.sdf_count * DMT_SDF_CLUSTER_FLOATS);
        std::memcpy(sdf_copy.data(), sdf->data.data() + k.sdf_off, sdf_copy.size() * 4u);
        d.sdf_clusters = sdf_copy.data();
        d.sdf_cluster_count = P.sdf_count;
    }
    std::vector<float> light_copy;
    if (lights) {
        light_copy.resize((std::size_t)P.light_count * kLightFloats);
        std::memcpy(light_copy.data(), lights->data.data() + k.lights_off, light_copy.size() * 4u);
        d.lights = light_copy.data();
    }
    dmt_build_params bp{};
    bp.flags = opt.device_bvh_build ? DMT_BUILD_DEVICE : 0u;
    dmt_scene* ns = nullptr;
    if (!check(dmt_scene_create(dev, &d, &bp, &ns), "dmt_scene_create")) return false;
    if (scene) dmt_scene_destroy(scene);   // waits for its frames
    scene = ns;
    scene_accum = nullptr;
    scene_prims = std::move(pr);
    if (tlas) scene_indices = tlas->indices;
    else scene_indices.clear();
    key = k;
    builds++;
    return true;
}

void HipDevice::dispatch(const HipCommandBuffer& cb) {
    try {
        dispatch_impl(cb);
    } catch (const std::bad_alloc&) {
        impl_->report("pathtrace: out of host memory");
    }
}

void HipDevice::dispatch_impl(const HipCommandBuffer& cb) {
    std::lock_guard lock(impl_->mu);
    Impl& I = *impl_;
    auto pl = I.pipelines.find(cb.pipeline_.id);
    if (pl == I.pipelines.end()) return;
    // other pipelines (clear, tonemap, bloom, overlays) are display passes the
    // software device also skips (SoftwareDevice.cpp:105-136)
    if (pl->second != "pathtrace") return;
    if (I.lost || !I.dev) {
        I.report("pathtrace: device lost");
        return;
    }
    PathtracePush P;
    std::string err;
    if (!DecodePathtracePush(cb.push_, cb.push_size_, I.opt.strict, &P, &err)) {
        I.report("pathtrace: " + err);
        return;
    }
    I.ignored = P.ignored;
    Texture* acc = I.texture(cb.textures_[slot::kAccumTexture]);
    Texture* out = I.texture(cb.textures_[slot::kOutputTexture]);
    if (!acc || !acc->dev) {
        I.report("pathtrace: storage-texture slot 1 holds no accumulation texture");
        return;
    }
    if (out && (out->w != acc->w || out->h != acc->h)) {
        I.report("pathtrace: output texture " + std::to_string(out->w) + "x" + std::to_string(out->h) +
                 " and accumulation texture " + std::to_string(acc->w) + "x" + std::to_string(acc->h) + " differ");
        return;
    }
    P.frame.width = acc->w;
    P.frame.height = acc->h;
    // the denoiser's textures, checked before any work so a refused dispatch changes nothing
    struct Aux {
        std::uint32_t slot;
        bool on;
        Texture* t;
    };
    Aux aux[] = {{slot::kDepthTexture, P.denoiser != 0u, nullptr},
                 {slot::kMotionTexture, P.denoiser != 0u, nullptr},
                 {slot::kNormalTexture, P.denoiser && P.write_normal, nullptr},
                 {slot::kAlbedoTexture, P.denoiser && P.write_albedo, nullptr},
                 {slot::kSpecularAlbedoTexture, P.denoiser && P.write_specular_albedo, nullptr},
                 {slot::kRoughnessTexture, P.denoiser && P.write_roughness, nullptr},
                 {slot::kSpecularDistanceTexture, P.denoiser && P.write_specular_distance, nullptr},
                 {slot::kCloudTransTexture, P.denoiser && P.clouds_mode == 0u, nullptr},
                 {slot::kDenoiseColorTexture, P.write_hdr_aux != 0u, nullptr}};
    for (Aux& a : aux) {
        if (!a.on) continue;
        a.t = I.texture(cb.textures_[a.slot]);
        if (!a.t || !a.t->dev) continue;   // not bound: not written
        if (a.t->w != acc->w || a.t->h != acc->h) {
            I.report("pathtrace: storage texture at slot " + std::to_string(a.slot) + " is " + std::to_string(a.t->w) +
                     "x" + std::to_string(a.t->h) + ", the frame " + std::to_string(acc->w) + "x" +
                     std::to_string(acc->h));
            return;
        }
    }
    if (!I.ensure_scene(cb, P)) return;
    if (I.scene_accum != acc->dev) {
        if (!I.check(dmt_scene_bind_accum(I.scene, acc->dev, acc->w, acc->h), "dmt_scene_bind_accum")) return;
        I.scene_accum = acc->dev;
    }
    auto written = [](Texture* t) {
        t->host_valid = false;
        t->gen++;
    };
    if (P.denoiser) {
        dmt_gbuffer_params gp{};
        std::memcpy(gp.curr_view_proj, P.curr_view_proj, sizeof(gp.curr_view_proj));
        std::memcpy(gp.prev_view_proj, P.prev_view_proj, sizeof(gp.prev_view_proj));
        gp.pixel_stride = 4;   // RGBA32F textures
        float** dst[] = {&gp.depth, &gp.motion, &gp.normal, &gp.albedo, &gp.specular_albedo, &gp.roughness,
                         &gp.specular_hit_distance, &gp.cloud_transmittance};
        for (int k = 0; k < 8; ++k)
            if (aux[k].on && aux[k].t && aux[k].t->dev) *dst[k] = aux[k].t->dev;
        if (!I.check(dmt_render_gbuffer(I.scene, &P.frame, &gp), "dmt_render_gbuffer")) return;
        for (int k = 0; k < 8; ++k)
            if (*dst[k]) written(aux[k].t);
    }
    // the moon disc of the gradient sky samples the moon texture at storage
    // slot 7 (PathTrace.slang:3306-3318)
    if (!P.env_map_present) {
        const std::uint64_t mid = cb.textures_[slot::kMoonTexture];
        Texture* mt = I.texture(mid);
        const std::uint64_t mgen = mt ? mt->gen : 0u;
        if (I.moon_scene != I.scene || I.moon_id != mid || I.moon_gen != mgen) {
            if (mt && mt->w && mt->h) {
                if (!I.host_copy(*mt)) return;
                if (!I.check(dmt_scene_set_moon_map(I.scene, mt->host.data(), mt->w, mt->h), "dmt_scene_set_moon_map"))
                    return;
            } else if (!I.check(dmt_scene_set_moon_map(I.scene, nullptr, 0, 0), "dmt_scene_set_moon_map")) {
                return;
            }
            I.moon_scene = I.scene;
            I.moon_id = mid;
            I.moon_gen = mgen;
        }
    }
    // the PBR atlas at storage slot 16 (RGBA8, the engine's 384-wide layout);
    // none bound or none uploaded: untextured
    {
        const std::uint64_t aid = cb.textures_[slot::kPbrAtlasTexture];
        Texture* at = I.texture(aid);
        const std::uint64_t agen = at ? at->gen : 0u;
        if (I.atlas_scene != I.scene || I.atlas_id != aid || I.atlas_gen != agen) {
            if (at && !at->rgba8.empty()) {
                if (!I.check(dmt_scene_set_pbr_atlas(I.scene, at->rgba8.data(), at->w, at->h),
                             "dmt_scene_set_pbr_atlas"))
                    return;
            } else if (!I.check(dmt_scene_set_pbr_atlas(I.scene, nullptr, 0, 0), "dmt_scene_set_pbr_atlas")) {
                return;
            }
            I.atlas_scene = I.scene;
            I.atlas_id = aid;
            I.atlas_gen = agen;
        }
    }
    if (!I.check(dmt_render(I.scene, &P.frame), "dmt_render")) return;
    written(acc);
    // exposure_state[0] scales the mean; invalid values read as 1
    // (SoftwareTracer.cpp:788-801)
    float exposure = 1.0f;
    if (const Buffer* eb = I.buffer(cb.buffers_[slot::kExposureBuffer]);
        eb && eb->data.size() >= cb.buffer_offsets_[slot::kExposureBuffer] + 4u)
        std::memcpy(&exposure, eb->data.data() + cb.buffer_offsets_[slot::kExposureBuffer], 4);
    dmt_tonemap_params tp{};
    tp.exposure = exposure;
    tp.op = P.tonemap_op;
    if (Texture* dc = aux[8].t; dc && dc->dev) {
        // the denoiser's colour input: raw HDR or tone-mapped, no OETF
        // (PathTrace.slang:7797-7811)
        tp.mode = P.hdr_pipeline ? DMT_TONEMAP_HDR : DMT_TONEMAP_LINEAR;
        if (!I.check(dmt_tonemap(I.scene, &tp, dc->dev), "dmt_tonemap(denoiser colour)")) return;
        written(dc);
    }
    if (out && out->dev) {
        tp.mode = DMT_TONEMAP_DISPLAY;
        if (!I.check(dmt_tonemap(I.scene, &tp, out->dev), "dmt_tonemap")) return;
        written(out);
    }
}

void HipDevice::clear_texture(TextureHandle h, const float rgba[4]) {
    std::lock_guard lock(impl_->mu);
    Impl& I = *impl_;
    Texture* t = I.texture(h.id);
    if (!t || !rgba) return;
    std::vector<float> v((std::size_t)t->w * t->h * 4u);
    for (std::size_t i = 0; i < v.size(); i += 4)
        for (int c = 0; c < 4; ++c) v[i + c] = rgba[c];
    if (!I.upload(*t, v.data())) return;
    t->host = std::move(v);
    t->host_valid = true;
    t->gen++;
}

void HipDevice::copy_buffer_to_texture(BufferHandle src, TextureHandle dst) {
    std::vector<std::uint8_t> bytes;
    {
        std::lock_guard lock(impl_->mu);
        const Buffer* b = impl_->buffer(src.id);
        if (!b) return;
        bytes = b->data;
    }
    (void)WriteTexture(dst, bytes.data(), bytes.size());
}

// ---- command buffer -------------------------------------------------------------

void HipCommandBuffer::BindComputePipeline(PipelineHandle p) { pipeline_ = p; }

void HipCommandBuffer::BindBuffer(std::uint32_t s, BufferHandle b, std::size_t offset) {
    if (s >= slot::kBufferSlots) return;
    buffers_[s] = b.id;
    buffer_offsets_[s] = offset;
}

void HipCommandBuffer::BindStorageTexture(std::uint32_t s, TextureHandle t) {
    if (s < slot::kTextureSlots) textures_[s] = t.id;
}

void HipCommandBuffer::BindAccelStruct(std::uint32_t s, AccelStructHandle a) {
    if (s < slot::kAccelSlots) accels_[s] = a.id;
}

void HipCommandBuffer::PushConstants(const void* data, std::size_t size) {
    if (!data) size = 0;
    size = std::min(size, sizeof(push_));
    if (size) std::memcpy(push_, data, size);
    push_size_ = size;
}

void HipCommandBuffer::Dispatch(std::uint32_t, std::uint32_t, std::uint32_t) { dev_->dispatch(*this); }

void HipCommandBuffer::CopyBufferToTexture(BufferHandle src, TextureHandle dst) {
    dev_->copy_buffer_to_texture(src, dst);
}

void HipCommandBuffer::ClearStorageTexture(TextureHandle t, const float rgba[4]) { dev_->clear_texture(t, rgba); }

void HipCommandBuffer::Barrier(const BarrierDesc&) {}   // one device stream orders everything

}  // namespace dmt::rhi
