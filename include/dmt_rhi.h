/*
 * dmt_rhi.h -- the HIP path tracer behind the reference engine's render
 * hardware interface (C++20, host side of libdmt_rhi.so over the C ABI in
 * dmt.h).
 *
 * dmt::rhi::HipDevice has the method set of pt::rhi::Device
 * (src/rhi/Device.h:24-434) and HipCommandBuffer that of pt::rhi::CommandBuffer
 * (src/rhi/CommandBuffer.h:15-44), with the same names, argument meanings and
 * error behaviour as the reference's software backend (SoftwareDevice,
 * src/rhi_software/SoftwareDevice.cpp): no exceptions, failures return a null
 * handle or false and leave a message in LastError(), a short push makes the
 * dispatch a no-op, IsDeviceLost() latches a device fault.  The types below
 * mirror src/rhi/Handles.h, Types.h, Resources.h and Swapchain.h field for
 * field, so a pt::rhi::Device subclass forwards to HipDevice call for call
 * (demont-engine_amd/csrc/rhi/pt_rhi_hip_device.cpp is that subclass).
 *
 * What a "pathtrace" dispatch does here (the engine's frame, Engine.cpp:
 * 6867-7217 binds, 7821-8676 PtPush, 9577-9582 dispatch):
 *   - the PtPush bytes are decoded into dmt_frame_params (DecodePathtracePush);
 *   - the analytic primitives (buffer slot 3, prim_count records of 112 B),
 *     the lights (slot 12, light_count records of 80 B), the TLAS (accel slot
 *     2, its instances flattened into one triangle mesh) and the env map
 *     (storage texture slot 5, RGBA32F) form a dmt_scene, built once and
 *     rebuilt only when one of them changes (a material-only change of the
 *     primitives goes through dmt_scene_set_materials instead, and a new TLAS
 *     over the same triangles -- moved instances -- through
 *     dmt_scene_update_mesh, a refit of the mesh's tree on the device);
 *   - the frame renders into the accum texture's device memory (slot 1, bound
 *     with dmt_scene_bind_accum, so the running mean survives a rebuild),
 *     with the push's depth of field, motion blur (v3 of the primitives as
 *     each sphere's previous centre), spectral dispersion and mesh translation;
 *   - the output texture (slot 0) receives sRGB(tonemap(mean * exposure))
 *     through dmt_tonemap, exposure = exposure_state[0] of buffer slot 6;
 *   - with denoiser_enabled, the bound G-buffer textures (depth 3, motion 4,
 *     normal 8, albedo 9, cloud transmittance 10, specular albedo 11,
 *     roughness 12, specular hit distance 13, each as its write flag says)
 *     receive dmt_render_gbuffer's outputs, and with write_hdr_aux the
 *     denoiser colour texture (slot 2) the frame's mean, raw or tone-mapped.
 * The light-tree nodes (slot 13) and env CDFs (slots 4/5) the engine uploads
 * are not read: the scene builds both from the lights and the env map with the
 * reference's own construction.  Work is stream-ordered on the device;
 * Submit is a no-op, WaitIdle / ReadbackTexture synchronise.
 */
#ifndef DMT_RHI_H
#define DMT_RHI_H

#include <cstddef>
#include <cstdint>
#include <memory>
#include <span>
#include <string>
#include <string_view>

#include "dmt.h"

namespace dmt::rhi {

/* src/rhi/Handles.h:12-16: 64-bit ids, 0 = null. */
struct BufferHandle      { std::uint64_t id = 0; explicit operator bool() const { return id != 0; } };
struct TextureHandle     { std::uint64_t id = 0; explicit operator bool() const { return id != 0; } };
struct PipelineHandle    { std::uint64_t id = 0; explicit operator bool() const { return id != 0; } };
struct AccelStructHandle { std::uint64_t id = 0; explicit operator bool() const { return id != 0; } };

/* src/rhi/Types.h:27-53 (same enumerator order and values). */
enum class TextureFormat : std::uint8_t { Unknown = 0, RGBA8_UNORM, RGBA8_SRGB, RGBA16F, RGBA32F, R32_UINT, R32F, RG16F, RG32F };
enum class TextureUsage : std::uint32_t { None = 0, Sampled = 1u << 0, Storage = 1u << 1, Present = 1u << 2 };
enum class BufferUsage : std::uint32_t {
    None = 0, Vertex = 1u << 0, Index = 1u << 1, Storage = 1u << 2, Uniform = 1u << 3,
    Indirect = 1u << 4, AccelInput = 1u << 5, Upload = 1u << 6,
};

/* src/rhi/Resources.h:15-71. */
struct BufferDesc {
    std::size_t size = 0;
    BufferUsage usage = BufferUsage::Storage;
    std::string_view debug_name;
};
struct TextureDesc {
    std::uint32_t width = 0;
    std::uint32_t height = 0;
    TextureFormat format = TextureFormat::RGBA8_UNORM;
    TextureUsage usage = TextureUsage::Sampled;
    std::string_view debug_name;
};
struct ComputePipelineDesc {
    std::string_view kernel_name;             /* "pathtrace" renders; other names are no-ops */
    std::span<const std::uint8_t> bytecode;   /* unused */
    std::string_view debug_name;
};
struct BLASDesc {
    const float* vertex_positions = nullptr;  /* float3 per vertex, copied */
    std::uint32_t vertex_count = 0;
    const std::uint32_t* indices = nullptr;   /* 3 per triangle, copied */
    std::uint32_t index_count = 0;
    std::string_view debug_name;
};
struct TLASInstance {
    AccelStructHandle blas;
    float transform[12]{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};   /* row-major 3x4 */
    std::uint32_t instance_id = 0;
    std::uint32_t mask = 0xFF;                 /* 0: never hit (rays query mask 0xFF) */
};
struct TLASDesc {
    std::span<const TLASInstance> instances;
    std::string_view debug_name;
};
struct BarrierDesc {
    enum class Stage : std::uint8_t { ComputeRead, ComputeWrite, Transfer, Present };
    Stage from = Stage::ComputeWrite;
    Stage to = Stage::ComputeRead;
};
/* src/rhi/Swapchain.h:9-14. */
struct FrameContext {
    TextureHandle swapchain_image;
    std::uint32_t width = 0;
    std::uint32_t height = 0;
    std::uint32_t frame_index = 0;
};

/* Bind slots of the engine's "pathtrace" pipeline (Engine.cpp:6867-7217;
 * storage textures and buffers are separate slot spaces). */
namespace slot {
constexpr std::uint32_t kOutputTexture = 0;    /* BindStorageTexture(0, swapchain image) */
constexpr std::uint32_t kAccumTexture = 1;     /* BindStorageTexture(1, accum)           */
constexpr std::uint32_t kDenoiseColorTexture = 2;   /* denoiser colour input (write_hdr_aux) */
constexpr std::uint32_t kDepthTexture = 3;     /* denoiser inputs (denoiser_enabled) ...  */
constexpr std::uint32_t kMotionTexture = 4;
constexpr std::uint32_t kEnvTexture = 5;       /* BindStorageTexture(5, env map)         */
constexpr std::uint32_t kMoonTexture = 7;      /* BindStorageTexture(7, moon_map)        */
constexpr std::uint32_t kNormalTexture = 8;    /* write_normal_gbuffer                   */
constexpr std::uint32_t kAlbedoTexture = 9;    /* write_albedo_gbuffer                   */
constexpr std::uint32_t kCloudTransTexture = 10;
constexpr std::uint32_t kSpecularAlbedoTexture = 11;
constexpr std::uint32_t kRoughnessTexture = 12;
constexpr std::uint32_t kSpecularDistanceTexture = 13;
constexpr std::uint32_t kPbrAtlasTexture = 16; /* BindStorageTexture(16, pbr_atlas), RGBA8 */
constexpr std::uint32_t kTlas = 2;             /* BindAccelStruct(2, scene TLAS)         */
constexpr std::uint32_t kPrimBuffer = 3;       /* BindBuffer(3, primitives)              */
constexpr std::uint32_t kEnvMarginalBuffer = 4;      /* BindBuffer(4, env_marginal_cdf)    */
constexpr std::uint32_t kEnvConditionalBuffer = 5;   /* BindBuffer(5, env_conditional_cdf) */
constexpr std::uint32_t kExposureBuffer = 6;   /* BindBuffer(6, exposure_state)          */
constexpr std::uint32_t kSdfBuffer = 10;       /* BindBuffer(10, sdf_clusters)           */
constexpr std::uint32_t kLightBuffer = 12;     /* BindBuffer(12, lights)                 */
constexpr std::uint32_t kMeshUvBuffer = 17;    /* BindBuffer(17, mesh_uvs)               */
constexpr std::uint32_t kBufferSlots = 24;     /* BindState sizes (SoftwareDevice.h:112-124) */
constexpr std::uint32_t kTextureSlots = 20;
constexpr std::uint32_t kAccelSlots = 4;
constexpr std::size_t kMaxPushBytes = 2048;    /* SoftwareDevice.h:148 */
}  // namespace slot

/* Byte offsets of the PtPush fields a pathtrace dispatch reads: the engine's
 * declaration (Engine.cpp:7243-7820; a plain sequence of 4-byte fields) and
 * its offset asserts (Engine.cpp:9388-9423: sun_and_mode 240, hosek_params
 * 1184, accum_params 784, tonemap_params 1600). */
namespace ptpush {
constexpr std::size_t kPosFovTan = 0, kFwdAspect = 16, kRightXyz = 32, kUpXyz = 48;
constexpr std::size_t kFrameIndex = 64, kResetAccum = 68, kMaxBounces = 72, kTlasPresent = 76;
constexpr std::size_t kPrimCount = 80, kSpp = 84, kDenoiserEnabled = 88, kEnvMapPresent = 92;
constexpr std::size_t kHaltonJitter = 96, kEnvIntensity = 104, kEnvTotalLuminance = 108;
constexpr std::size_t kCurrViewProj = 112, kPrevViewProj = 176;
constexpr std::size_t kHeadBytes = 112;        /* PtPushHead (SoftwareTracer.cpp:38-56) */
constexpr std::size_t kSunAndMode = 240, kW2jRow0 = 272, kW2jRow1 = 288, kW2jRow2 = 304;
constexpr std::size_t kDofParams = 320, kDofBokehParams = 336, kSpectralParams = 352, kVolParams = 368;
constexpr std::size_t kCloudsP1 = 384, kCloudsP3 = 416, kMoonDirPhase = 448, kMoonExtra = 464, kSunExtra = 480;
constexpr std::size_t kHdriLightsDir = 496, kHdriLightsCol = 624;
constexpr std::size_t kHdriLightsCount = 752, kWriteNormalGbuffer = 756, kWriteAlbedoGbuffer = 760;
constexpr std::size_t kWriteHdrAux = 764, kMisEnabled = 768, kAccumParams = 784, kSdfParams = 816, kSdfParamsF = 832;
constexpr std::size_t kWriteShadowVis = 856, kLightCount = 880, kLightTreeEnabled = 888;
constexpr std::size_t kWriteSpecularAlbedo = 896, kWriteRoughness = 900, kWriteSpecularHitDistance = 904;
constexpr std::size_t kWaterParams0 = 912, kWaterParams1 = 928, kWaterState = 1040, kOceanParams0 = 1120;
constexpr std::size_t kRestirEnabled = 944, kSmokeParams = 960, kMotionBlurParams = 1024, kCloudsRuntime = 1072;
constexpr std::size_t kMeshMotionPrev = 1088, kMeshMotionCurr = 1104;
constexpr std::size_t kMeshTexIndices = 1152, kMatOverrideCount = 1200, kTonemapParams = 1600, kSunExtra2 = 1616;
constexpr std::size_t kSize = 1792;            /* sizeof(PtPush) with kMaxHdriLights = kMaxMatOverrides = 8 */
}  // namespace ptpush

/* One pathtrace push, decoded. */
struct PathtracePush {
    dmt_frame_params frame{};       /* width/height 0: they come from the bound textures */
    std::uint32_t tlas_present = 0;
    std::uint32_t prim_count = 0;
    std::uint32_t light_count = 0;
    std::uint32_t env_map_present = 0;
    float env_intensity = 1.0f;
    float env_total_luminance = 0.0f;   /* the engine's CDF total (used with its CDF buffers) */
    std::uint32_t sdf_count = 0;    /* sdf_params.x: SDF clusters at buffer slot 10 */
    std::uint32_t tonemap_op = 0;   /* r_tonemap_op; values above 4 select ACES */
    std::uint32_t motion_blur = 0;  /* motion_blur_params.x > 0.5 (also frame.motion_blur) */
    /* denoiser inputs (PathTrace.slang:5463-5648, 7797-7811) */
    std::uint32_t denoiser = 0;               /* denoiser_enabled: depth + motion */
    std::uint32_t write_normal = 0, write_albedo = 0, write_specular_albedo = 0;
    std::uint32_t write_roughness = 0, write_specular_distance = 0;
    std::uint32_t write_hdr_aux = 0;          /* the denoiser's colour input */
    std::uint32_t hdr_pipeline = 0;           /* w2j_row0.w > 0.5: that input raw, not tone-mapped */
    std::uint32_t clouds_mode = 0;            /* clouds_runtime.x (0: cloud transmittance written) */
    float curr_view_proj[16]{};
    float prev_view_proj[16]{};
    std::string ignored;            /* lenient mode: features asked for and not rendered */
};

/* Decodes the PtPush bytes of a pathtrace dispatch.  Returns false, with the
 * reason in *error, for a push shorter than the 112-byte head (the software
 * tracer's no-op, SoftwareTracer.cpp:686-690) and, when strict, for a push that
 * turns on a feature this backend does not render: a procedural / Hosek sky
 * (sun_and_mode.w 2 or 3), participating media, the SDF iteration heat map, ReSTIR,
 * material overrides, smoke, the SIGMA shadow-visibility output, the FFT
 * ocean (ocean_params0.x), or a per-frame denoiser colour input without
 * reset_accum.
 * Lenient decoding renders without them and lists them in out->ignored.  The
 * gradient sky's sun and moon (sun_and_mode, sun_extra, sun_extra2,
 * moon_dir_phase, moon_extra), the HDRI cluster lights (hdri_lights_*) and,
 * while a denoiser runs, the Halton jitter are decoded into the frame, and
 * the SDF clusters' count and sphere-trace parameters (sdf_params.xy,
 * sdf_params_f.x) into sdf_count and the frame, and the mesh's PBR map tiles
 * (mesh_tex_indices) into frame.mesh_tex with DMT_FRAME_MESH_TEXTURES when
 * any is set, and the water (water_params0, water_params1, water_state) into
 * the frame's water fields.
 * Fields past the pushed size read as zero. */
bool DecodePathtracePush(const void* data, std::size_t size, bool strict, PathtracePush* out, std::string* error);

struct HipDeviceOptions {
    std::int32_t device_index = 0;
    std::uint32_t width = 1280;     /* initial swapchain size (the software device's 1280x720 sentinel) */
    std::uint32_t height = 720;
    bool strict = true;             /* see DecodePathtracePush */
    bool device_bvh_build = false;  /* mesh trees built on the device (DMT_BUILD_DEVICE): faster
                                       rebuilds after mesh edits, somewhat slower traversal */
};

class HipDevice;

/* pt::rhi::CommandBuffer (CommandBuffer.h:15-44): records binds and push
 * constants; Dispatch runs the bound pipeline at once (stream-ordered). */
class HipCommandBuffer {
public:
    void BindComputePipeline(PipelineHandle p);
    void BindBuffer(std::uint32_t slot, BufferHandle b, std::size_t offset = 0);
    void BindStorageTexture(std::uint32_t slot, TextureHandle t);
    void BindAccelStruct(std::uint32_t slot, AccelStructHandle a);
    void PushConstants(const void* data, std::size_t size);   /* truncated to 2048 B */
    void Dispatch(std::uint32_t gx, std::uint32_t gy, std::uint32_t gz);
    void CopyBufferToTexture(BufferHandle src, TextureHandle dst);
    void ClearStorageTexture(TextureHandle t, const float rgba[4]);
    void Barrier(const BarrierDesc& d);

private:
    friend class HipDevice;
    explicit HipCommandBuffer(HipDevice* d) : dev_(d) {}
    HipDevice* dev_;
    PipelineHandle pipeline_{};
    std::uint64_t buffers_[slot::kBufferSlots]{};
    std::size_t buffer_offsets_[slot::kBufferSlots]{};
    std::uint64_t textures_[slot::kTextureSlots]{};
    std::uint64_t accels_[slot::kAccelSlots]{};
    std::uint8_t push_[slot::kMaxPushBytes]{};
    std::size_t push_size_ = 0;
};

/* pt::rhi::Device (Device.h:24-434) for the HIP path tracer. */
class HipDevice {
public:
    explicit HipDevice(const HipDeviceOptions& options = {});
    ~HipDevice();
    HipDevice(const HipDevice&) = delete;
    HipDevice& operator=(const HipDevice&) = delete;

    BufferHandle CreateBuffer(const BufferDesc& d);
    TextureHandle CreateTexture(const TextureDesc& d);
    PipelineHandle CreateComputePipeline(const ComputePipelineDesc& d);
    AccelStructHandle CreateBLAS(const BLASDesc& d);
    AccelStructHandle CreateTLAS(const TLASDesc& d);
    void DestroyBuffer(BufferHandle h);
    void DestroyTexture(TextureHandle h);
    void DestroyPipeline(PipelineHandle h);
    void DestroyAccelStruct(AccelStructHandle h);

    void WriteBuffer(BufferHandle h, const void* src, std::size_t size, std::size_t dst_offset = 0);
    bool WriteTexture(TextureHandle h, const void* src, std::size_t src_size);

    FrameContext BeginFrame();
    void EndFrame(HipCommandBuffer* cb);
    HipCommandBuffer* AcquireCommandBuffer();
    void Submit(HipCommandBuffer* cb);
    void WaitIdle();
    void Resize(int w, int h);

    bool ReadbackTexture(TextureHandle h, void* dst, std::size_t dst_size, std::uint32_t* out_w,
                         std::uint32_t* out_h);
    bool ReadbackBuffer(BufferHandle h, void* dst, std::size_t bytes);

    bool SupportsHardwareRT() const { return true; }   /* the engine takes the BLAS/TLAS route */
    const char* DeviceName() const;
    std::size_t CurrentAllocatedBytes() const;
    bool IsDeviceLost() const;

    /* Beyond the reference interface. */
    const std::string& LastError() const;          /* last failure ("" if none) */
    dmt_scene* ActiveScene() const;                /* scene of the last pathtrace dispatch */
    std::uint64_t SceneBuilds() const;             /* dmt_scene_create calls so far */
    std::uint64_t MaterialUpdates() const;         /* dmt_scene_set_materials calls so far */
    std::uint64_t MeshRefits() const;              /* dmt_scene_update_mesh calls so far */
    const std::string& IgnoredFeatures() const;    /* lenient mode: of the last dispatch */

private:
    friend class HipCommandBuffer;
    void dispatch(const HipCommandBuffer& cb);
    void dispatch_impl(const HipCommandBuffer& cb);
    void clear_texture(TextureHandle t, const float rgba[4]);
    void copy_buffer_to_texture(BufferHandle src, TextureHandle dst);
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

/* CreateSoftwareDevice's counterpart (SoftwareDevice.cpp:1165); null when no
 * HIP device can be opened (the reason is on stderr). */
std::unique_ptr<HipDevice> CreateHipDevice(const HipDeviceOptions& options = {});

}  // namespace dmt::rhi

#endif /* DMT_RHI_H */
