/*
 * dmt.h -- C ABI of the MI355X (gfx950) wavefront path tracer.
 *
 * This is the drop-in boundary for the reference's software path-trace
 * backend.  In the reference the engine reaches the tracer through the RHI:
 *   Device::Create(BackendType, ...)                  src/rhi/Device.cpp:20-48
 *   CreateBLAS / CreateTLAS (Embree scene build)      src/rhi_software/SoftwareDevice.cpp:337-421
 *   WriteBuffer (prims, lights, light tree, env CDF)  src/rhi_software/SoftwareDevice.cpp:539-558,
 *                                                     src/engine/Engine.cpp:4743-5002, 5235-5600
 *   PushConstants(PtPush) + Dispatch("pathtrace")     src/engine/Engine.cpp:9577-9582,
 *                                                     src/rhi_software/SoftwareDevice.cpp:105-136
 *   RunPathTraceKernel(...)                           src/rhi_software/SoftwareTracer.h:40-41
 *   ReadbackTexture(accum)                            src/rhi_software/SoftwareDevice.cpp:458-483
 * The entry points below replace those calls one for one (each declaration
 * names the reference interface it replaces).  Plain pointers and sizes only;
 * every host buffer is copied on entry; no callbacks.  One thread per device.
 * Work is ordered on the device's HIP stream; functions that return host data
 * synchronise that stream first.
 *
 * Wire formats are the reference's:
 *   analytic primitive  7 float4 = 112 B (Engine.cpp:4924-4994):
 *     v0 = centre.xyz, radius  (plane: normal.xyz, d with n.p + d = 0)
 *     v1 = albedo.rgb, roughness
 *     v2 = type (0 sphere, 1 plane), material (0 Lambert, 1 metal, 2 dielectric),
 *          ior, user id (uint32 bits)
 *     v3 = previous centre (the motion-blur shutter lerps from it to v0)
 *     v4 = emission.rgb, 0
 *     v5 = orientation quaternion xyzw (planes: normal = quatRotate(q, v0.xyz))
 *     v6 = texture tiles (ignored: no PBR atlas)
 *   analytic light      5 float4 = 80 B (Engine.cpp:5235-5300):
 *     v0 = pos.xyz, radius; v1 = intensity.rgb, type (0 point, 1 spot, 2 sphere, 3 quad)
 *     v2 = dir.xyz (spot axis / quad normal), cos_outer
 *     v3 = quad u half-extent vector.xyz (spot: x = cos_inner), v_half
 *     v4 = orientation quaternion xyzw
 *   mesh                float[3V] positions, uint32[3T] indices
 *   env map             float[3*W*H] linear RGB, row 0 = +Y pole (lat-long)
 *   accumulation        float4 per pixel: running-mean radiance rgb, sample count
 */
#ifndef DMT_H
#define DMT_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define DMT_ABI_VERSION 8u

typedef int32_t dmt_status;
#define DMT_OK                  0
#define DMT_ERR_INVALID_ARG    -1   /* bad pointer, size or enum value       */
#define DMT_ERR_OUT_OF_MEMORY  -2   /* device or host allocation failed       */
#define DMT_ERR_NO_DEVICE      -3   /* no HIP device / device index invalid   */
#define DMT_ERR_DEVICE         -4   /* a HIP call failed                      */
#define DMT_ERR_UNSUPPORTED    -5   /* feature outside this backend's scope   */
#define DMT_ERR_DEVICE_LOST    -6   /* device reported a fault; terminal      */
#define DMT_ERR_LIMIT          -7   /* size beyond a documented maximum       */

#define DMT_PRIM_FLOATS   28u
#define DMT_LIGHT_FLOATS  20u
#define DMT_LTREE_FLOATS  16u
#define DMT_MAX_SPP       32u       /* r_spp clamp (Engine.cpp:7223-7228) */
#define DMT_MAX_BOUNCES   64u

typedef struct dmt_device dmt_device;
typedef struct dmt_scene  dmt_scene;

typedef struct dmt_scene_desc {
    const float*    prims;      uint32_t prim_count;
    const float*    lights;     uint32_t light_count;
    const float*    positions;  uint32_t vertex_count;
    const uint32_t* indices;    uint32_t index_count;
    const float*    env_rgb;    uint32_t env_width;  uint32_t env_height;
    float           env_intensity;
    uint32_t        reserved;
    /* Optional importance CDF of the env map, as the engine uploads it
     * (env_marginal_cdf H floats, env_conditional_cdf W*H floats, and
     * env_total_luminance; Engine.cpp:4571-4605).  NULL: the scene builds it
     * from env_rgb (BuildEnvCdf, HdrImage.cpp:169-219).  An engine that
     * extracts HDRI cluster lights builds it with those clusters' pixels
     * left out (dmt_build_env_cdf_masked), which the path tracer reads as
     * the clusters' directions (isExtractedClusterDir). */
    const float*    env_marginal;
    const float*    env_conditional;
    double          env_total_luminance;
    /* Optional SDF clusters (the engine's sdf_clusters buffer, slot 10;
     * Engine.cpp:5076-5190): DMT_SDF_CLUSTER_FLOATS floats each --
     * hdr0 (aabb_min.xyz, node_count bits), hdr1 (aabb_max.xyz, material
     * bits), hdr2 (albedo, roughness), hdr3 (ior, 0, 0, 0), then 8 nodes of
     * (op, shape, child_a, child_b bits), params, (center.xyz, 0) in
     * post-order.  Sphere-traced as PathTrace.slang:2163-2306 does in the
     * build without procedural ops; fractal roots are DMT_ERR_UNSUPPORTED. */
    const float*    sdf_clusters;
    uint32_t        sdf_cluster_count;
    uint32_t        reserved2;
    /* Optional per-vertex mesh UVs, 2 floats per vertex (the engine's
     * mesh_uvs buffer, slot 17, parallel to the positions): the mesh's PBR
     * maps (dmt_frame_params.mesh_tex) are sampled at the hit's barycentric
     * UV.  NULL: the mesh is untextured. */
    const float*    mesh_uvs;
} dmt_scene_desc;

#define DMT_SDF_CLUSTER_FLOATS 112u

/* Per-frame parameters: the PtPush fields the path reads
 * (Engine.cpp:7243, 7821-8676; camera basis Engine.cpp:7821-7828). */
typedef struct dmt_frame_params {
    float    pos_fovtan[4];     /* camera position, tan(fov_y / 2)           */
    float    fwd_aspect[4];     /* forward, aspect = width / height          */
    float    right_xyz[4];
    float    up_xyz[4];
    uint32_t width, height;
    uint32_t frame_index;       /* seeds pcgHash(x*1973 + y*9277 + f*26699) */
    uint32_t spp;               /* 1..32                                     */
    uint32_t max_bounces;       /* r_max_bounces                             */
    uint32_t reset_accum;
    uint32_t mis_enabled;       /* r_mis                                     */
    uint32_t light_tree_enabled;/* r_light_tree                              */
    uint32_t refract_bounces;   /* r_refract_bounces (transmittance budget)  */
    float    firefly_clamp;     /* r_firefly_clamp, <= 0 disables            */
    float    ema_alpha;         /* r_accum_ema_alpha                         */
    uint32_t flags;             /* DMT_FRAME_* bits; 0 = reference behaviour */
    /* Camera-sample features, all off when zero.  Per sample, after the
     * jitter, the reference draws the hero wavelength (spectral), then the
     * shutter time (motion_blur), then the lens point (dof[0] > 0)
     * (PathTrace.slang:5729-5816).  A frame that uses any of them renders
     * in one launch of the frame-completion kernel, which starts the camera
     * paths itself. */
    float    dof[4];            /* dof_params: aperture radius (0 = pinhole),
                                   focal distance, blade count (< 3 = round),
                                   blade rotation in radians               */
    float    cat_eye;           /* dof_bokeh_params.x, 0..1                  */
    uint32_t motion_blur;       /* motion_blur_params.x: a shutter time t per
                                   sample; spheres at lerp(prev_pos, pos, t) */
    uint32_t spectral;          /* spectral_params.x: hero wavelength per
                                   sample, dielectric index by Cauchy's law  */
    float    cauchy_c;          /* spectral_params.y, Cauchy C in um^2       */
    float    mesh_shift_prev[3];/* mesh_motion_prev.xyz                      */
    uint32_t mesh_shift_active; /* mesh_motion_prev.w > 0.5: the mesh is drawn
                                   translated by lerp(prev, curr, t), t = 1
                                   without motion blur (PathTrace.slang:2483) */
    float    mesh_shift_curr[3];/* mesh_motion_curr.xyz                      */
    uint32_t reserved;          /* 0                                         */
    /* Sample 0's sub-pixel offset when flags has DMT_FRAME_FIXED_JITTER: the
     * engine's per-frame Halton(2,3) offset while a denoiser runs
     * (halton_jitter, PathTrace.slang:5730-5731, Engine.cpp:8245-8254),
     * each in [-0.5, 0.5]; that sample draws no jitter. */
    float    jitter[2];
    /* Sky lights, all zero for none.  Without an env map the gradient sky
     * carries the sun and moon: their discs in the sky (sunDisc, moonDisc,
     * PathTrace.slang:3112, 3212) and, when above the horizon, their NEE at
     * Lambert vertices (:6546-6667).  With an env map, hdri_light_count
     * directional lights extracted from its bright clusters get their own NEE
     * (:6493-6518, :3749). */
    float    sun[4];            /* sun_and_mode.xyz: unit direction; w: 0    */
    float    sun_extra[4];      /* r_sun_size, distance ratio, vertical
                                   flattening (1 = round), 0               */
    float    sun_disc_brightness; /* sun_extra2.x                            */
    float    moon[4];           /* moon_dir_phase: unit direction, phase (rad) */
    float    moon_extra[4];     /* r_moon_size, distance ratio, disc
                                   brightness, 0                            */
    uint32_t hdri_light_count;  /* hdri_lights_count, <= 8                   */
    float    hdri_dir[8][4];    /* hdri_lights_dir: unit direction, pick
                                   probability (summing to 1)              */
    float    hdri_col[8][4];    /* hdri_lights_col: rgb flux, 0              */
    /* SDF sphere trace (sdf_params.y, sdf_params_f.x), read when the scene
     * has SDF clusters: the step cap, 1..256 (r_sdf_max_iters, default 128),
     * and the surface distance, > 0 (r_sdf_epsilon, default 1e-4; the
     * engine clamps it to >= 1e-7). */
    uint32_t sdf_max_iters;
    float    sdf_epsilon;
    /* The mesh's PBR map tiles (mesh_tex_indices: albedo, normal, roughness,
     * metallic; 0xFFFFFFFF = none), read when flags has
     * DMT_FRAME_MESH_TEXTURES and the scene has mesh UVs and an atlas. */
    uint32_t mesh_tex[4];
    /* Water (water_params0, water_params1, water_state; PathTrace.slang:
     * 5089-5296, 5847-5872, 6014-6050, 6768-6906): surfaces of material 3
     * (DMT_MAT_WATER) bend their normal by drifting value-noise waves,
     * reflect with Schlick's F0 0.02 or refract into and out of the water
     * medium, in which a ray loses exp(-water_absorb * distance) per segment
     * and, escaping, returns deep_color.  Frames with water render in the
     * water variants of the frame-completion and batched kernels; the
     * adjoint refuses them. */
    float    water_absorb[3];   /* water_params0.xyz, 1/m, >= 0            */
    float    water_ior;         /* water_params0.w: index of surfaces whose own is <= 0 */
    float    wave[4];           /* water_params1: scale (cycles/m), amplitude, speed, time (s) */
    uint32_t camera_in_water;   /* water_state.x > 0.5: camera paths start in the water */
    float    deep_color[3];     /* water_state.yzw, >= 0                   */
} dmt_frame_params;

/* Russian roulette off (the reference applies it from bounce 2 on,
 * PathTrace.slang:6908-6913).  Its kill test compares a random number with
 * a throughput-dependent probability, a decision that flips under parameter
 * perturbations; differentiable renders checked against finite differences
 * turn it off. */
#define DMT_FRAME_NO_RUSSIAN_ROULETTE 1u
/* Sample 0 takes jitter[] instead of two random draws (the engine sets it
 * while a denoiser runs, PathTrace.slang:5730). */
#define DMT_FRAME_FIXED_JITTER 2u
/* mesh_tex holds the mesh's map tiles (the engine sets mesh_tex_indices every
 * frame; a zero-initialised frame keeps the mesh untextured). */
#define DMT_FRAME_MESH_TEXTURES 4u
/* Material id of water surfaces (MAT_WATER, PathTrace.slang:1403). */
#define DMT_MAT_WATER 3u

typedef struct dmt_build_params {
    uint32_t sah_bins;          /* binned-SAH bin count (0 = default 16)     */
    uint32_t max_leaf;          /* triangles per leaf (0 = default 4, <= 8)  */
    uint32_t flags;             /* DMT_BUILD_* bits                          */
    uint32_t reserved;
} dmt_build_params;
/* Build the triangle BVH on the device: a linear BVH over the Morton order of
 * the triangle centroids (Karras 2012), collapsed to 4-wide nodes -- for
 * meshes whose triangles change, where the host's binned-SAH build is the
 * slow step (no counterpart in the reference, whose builds run on the host:
 * TriangleBvh::Build, src/renderer/TriangleBvh.cpp:33-169; Embree scenes,
 * SoftwareDevice.cpp:337-421).  Same images (the closest hit does not depend
 * on the tree); the tree traverses somewhat slower than the SAH one. */
#define DMT_BUILD_DEVICE 1u

typedef struct dmt_hit {
    float    t;
    float    normal[3];
    uint32_t hit;               /* 0 miss, 1 analytic, 2 mesh, 3 SDF cluster */
    uint32_t prim;              /* analytic: input index; mesh: triangle id;
                                   SDF: cluster index                        */
    uint32_t mat;
    uint32_t user_id;           /* analytic v2.w bits; 0 for mesh            */
} dmt_hit;

typedef struct dmt_stats {
    uint64_t primary_rays;      /* camera rays traced                        */
    uint64_t extension_rays;    /* bounce rays traced (b >= 1)               */
    uint64_t shadow_rays;       /* traceScene steps inside transmittance()   */
    uint64_t frames;
    uint64_t extend_launches;   /* closest-hit kernel launches               */
    double   extend_ms;         /* HIP-event time summed over those launches */
    uint64_t node_visits;       /* filled by dmt_scene_measure_traffic:      */
    uint64_t tri_tests;         /*   all traversals (extend + shadow)        */
    uint64_t sphere_tests;
    uint64_t ext_node_visits;   /*   closest-hit (extend) kernel only        */
    uint64_t ext_tri_tests;
    uint64_t ext_sphere_tests;
    uint64_t scene_bytes;       /* device bytes of the acceleration data     */
    uint32_t bvh_nodes;
    uint32_t bvh_depth;
    uint64_t tail_node_visits;  /*   frame-completion (tail) kernel: its     */
    uint64_t tail_tri_tests;    /*   extension and shadow traversals         */
    uint64_t tail_sphere_tests;
    uint64_t ext_rays;          /*   rays of the closest-hit kernel          */
    uint64_t tail_launches;     /* frame-completion kernel launches (timed   */
    double   tail_ms;           /*   like extend_launches / extend_ms)       */
} dmt_stats;

/* ABI version check. */
dmt_status dmt_get_abi_version(uint32_t* out_version);

/* Replaces Device::Create(BackendType::Software, ...) / CreateSoftwareDevice
 * (src/rhi/Device.cpp:20-48, src/rhi_software/SoftwareDevice.cpp:1165). */
dmt_status dmt_device_create(int32_t device_index, dmt_device** out_device);
/* Destroying a device that still has scenes defers the teardown until the
 * last of its scenes is destroyed. */
void       dmt_device_destroy(dmt_device* dev);
/* hipStream_t the device orders its work on (for caller-side events). */
dmt_status dmt_device_stream(dmt_device* dev, void** out_stream);
/* Replaces Device::WaitIdle (src/rhi_software/SoftwareDevice.h:181). */
dmt_status dmt_device_synchronize(dmt_device* dev);
/* Replaces Device::IsDeviceLost (src/rhi/Device.h:157-164). */
int32_t    dmt_device_is_lost(const dmt_device* dev);

/* Replaces CreateBLAS/CreateTLAS + the engine's prim/light/light-tree/env
 * uploads (SoftwareDevice.cpp:337-421, Engine.cpp:3752-3900, 4743-5002,
 * 5235-5600, 4571-4605): copies the scene, builds the binned-SAH BVH4, the
 * light tree and the env CDF on the host, and uploads them to HBM. */
dmt_status dmt_scene_create(dmt_device* dev, const dmt_scene_desc* desc,
                            const dmt_build_params* build, dmt_scene** out_scene);
void       dmt_scene_destroy(dmt_scene* scene);

/* Replaces PushConstants(PtPush) + Dispatch("pathtrace") ->
 * RunPathTraceKernel (SoftwareTracer.cpp:684-1015, PathTrace.slang:5450-7781):
 * renders one frame into the scene's accumulation buffer (device resident),
 * stream-ordered, no host synchronisation. */
dmt_status dmt_render(dmt_scene* scene, const dmt_frame_params* frame);

/* Renders frame_count consecutive frames (frame_index, frame_index + 1, ...;
 * reset_accum applies to the first) into the accumulation buffer: the
 * engine's progressive loop over Dispatch("pathtrace") (Engine.cpp:12742-12945
 * -> :9582) issued as one call.  Identical to frame_count dmt_render calls. */
dmt_status dmt_render_frames(dmt_scene* scene, const dmt_frame_params* first_frame, uint32_t frame_count);

/* Replaces ReadbackTexture(accum_hdr) (SoftwareDevice.cpp:458-483):
 * full-or-nothing copy of width*height float4 into rgba. */
dmt_status dmt_readback_accum(dmt_scene* scene, float* rgba, size_t float_count);
/* Device address of the accumulation buffer (width*height float4). */
dmt_status dmt_accum_device_ptr(dmt_scene* scene, void** out_ptr);

/* Caller-owned accumulation memory, for an engine whose accumulation texture
 * outlives scene rebuilds: the reference engine binds one accum texture at
 * storage-texture slot 1 for every frame (Engine.cpp:6867-6870) and its
 * contents, not the scene, carry the running mean.  d_rgba is width*height
 * float4 of device memory on the scene's device, 16-byte aligned; its
 * contents are the history (nothing is cleared), and every frame rendered
 * must have this width and height.  NULL returns the scene to a buffer of its
 * own (allocated and zeroed at the next render).  Waits for the frames in
 * flight on the device.  Not allowed while the scene belongs to a batch. */
dmt_status dmt_scene_bind_accum(dmt_scene* scene, void* d_rgba, uint32_t width, uint32_t height);

/* The moon texture the moon disc samples (moon_map, Engine.cpp:4612-4653,
 * bound at storage-texture slot 7): width*height RGBA floats, row 0 the
 * disc's north; copied.  NULL (or never set): the disc's surface reads 0.
 * Waits for the scene's frames in flight. */
dmt_status dmt_scene_set_moon_map(dmt_scene* scene, const float* rgba, uint32_t width, uint32_t height);

/* The PBR material atlas (pbr_atlas, Engine.cpp:3984-4148, bound at
 * storage-texture slot 16): RGBA8, width 384 (a 256-square base image per tile
 * and its mips 1..8 in the right column, Engine::BuildPbrTileMips), height a
 * multiple of 256 (one 256-row band per tile); copied.  A texel reads as
 * byte * (1/255).  An analytic primitive's v6 holds its albedo, normal,
 * roughness and metallic tiles (bit patterns; 0xFFFFFFFF = none), the mesh's
 * come with the frame; a tile applies when it is below height / 256
 * (applyPbrTextures, PathTrace.slang:1435-1605).  Frames of a scene with
 * textured surfaces render in the PBR variants of the frame-completion and
 * batched kernels; the adjoint refuses them.  NULL: no atlas.  Waits for the
 * scene's frames in flight. */
dmt_status dmt_scene_set_pbr_atlas(dmt_scene* scene, const uint8_t* rgba8, uint32_t width, uint32_t height);

/* Display output (r_tonemap_op, exposure_state[0]). */
#define DMT_TONEMAP_DISPLAY  0u   /* sRGB OETF of tonemap(mean * exposure)     */
#define DMT_TONEMAP_LINEAR   1u   /* tonemap(max(mean, 0) * exposure), no OETF */
#define DMT_TONEMAP_HDR      2u   /* max(mean, 0): the raw HDR image           */
typedef struct dmt_tonemap_params {
    float    exposure;          /* scales the mean; <= 0 or non-finite -> 1  */
    uint32_t op;                /* 0 ACES, 1 AgX, 2 Khronos PBR Neutral,
                                   3 Reinhard, 4 linear clamp; others -> 0   */
    uint32_t mode;              /* DMT_TONEMAP_*                              */
    uint32_t reserved;
} dmt_tonemap_params;
/* DMT_TONEMAP_DISPLAY replaces the output stage of RunPathTraceKernel
 * (SoftwareTracer.cpp:788-801, 995-1007), which writes the slot-0 output
 * texture; LINEAR and HDR are the denoiser's colour input of the write_hdr_aux
 * path (PathTrace.slang:7797-7811, r_hdr_pipeline off / on) when each frame
 * resets the accumulation (the engine resets every frame while a denoiser
 * runs), so the mean is that frame's radiance.  rgb as the mode says,
 * alpha = 1, as width*height float4 into device memory d_out_rgba.
 * Stream-ordered after the scene's last frame; later frames wait for it
 * before they update the accumulation buffer. */
dmt_status dmt_tonemap(dmt_scene* scene, const dmt_tonemap_params* params, void* d_out_rgba);

/* Denoiser inputs of a frame (PathTrace.slang:5463-5648, written when the
 * engine sets denoiser_enabled): a closest-hit query along every pixel-centre
 * camera ray (no jitter) and, per pixel, into device memory (NULL: not
 * written; `pixel_stride` floats from one pixel to the next, >= the floats an
 * output writes, 4 for RGBA32F textures):
 *   depth                 1  view depth t * dot(d, normalize(forward)); 1e10 on a miss
 *   motion                2  previous - current pixel position of the hit point
 *                            (of the point 1000 along the ray on a miss)
 *   normal, albedo        4  the hit's normal / albedo, 0; zero on a miss
 *   specular_albedo       4  F0: metal albedo, dielectric 0.04, Lambert 0; 0
 *   roughness             1  saturate(roughness); 1 for Lambert and misses
 *   specular_hit_distance 1  t * (1 - saturate(roughness)); 0 on a miss
 *   cloud_transmittance   1  1 (no volumetric clouds in this backend)
 * Pixel positions are ((x / w) * 0.5 + 0.5) * width, ((-y / w) * 0.5 + 0.5) *
 * height of clip = M * (p, 1), M column-major as the engine packs
 * curr_view_proj / prev_view_proj (Engine.cpp:8257-8259).  Stream-ordered on
 * the device stream. */
typedef struct dmt_gbuffer_params {
    float    curr_view_proj[16];
    float    prev_view_proj[16];
    uint32_t pixel_stride;
    uint32_t reserved;
    float*   depth;
    float*   motion;
    float*   normal;
    float*   albedo;
    float*   specular_albedo;
    float*   roughness;
    float*   specular_hit_distance;
    float*   cloud_transmittance;
} dmt_gbuffer_params;
dmt_status dmt_render_gbuffer(dmt_scene* scene, const dmt_frame_params* frame, const dmt_gbuffer_params* params);

/* traceScene() closest-hit query (PathTrace.slang:2331-2693) for n rays,
 * 8 floats per ray (ox, oy, oz, -, dx, dy, dz, -); host in, host out.
 * SDF clusters are traced with the engine's default 128 steps and 1e-4. */
dmt_status dmt_trace_closest(dmt_scene* scene, const float* rays, uint32_t n, dmt_hit* out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): dmt_trace_closest on caller-owned device memory, with no
 * hipMalloc, hipFree or other allocating call.  d_rays holds n rays of 8
 * floats (ox, oy, oz, -, dx, dy, dz, -), 16-byte aligned; d_out receives n
 * dmt_hit records, as dmt_trace_closest returns them.  Stream-ordered on the
 * device stream (dmt_device_stream), no host synchronisation: d_rays must be
 * complete when the call is made and d_out is valid after that stream's
 * later work.  Its traversal-stack scratch is a static device array of each
 * GPU (24 MiB, held until the process ends): the scene's trees must be at
 * most (DMT_NOALLOC_SPILL_DEPTH + 12) / 3 = 20 levels deep, deeper ones are
 * refused with DMT_ERR_LIMIT; any number of rays.  Calls on one GPU share
 * that array, so calls that overlap on it (from several threads, or from two
 * dmt_device handles of the same GPU, whose streams differ) race. */
#define DMT_NOALLOC_BLOCKS       512u   /* workgroups of the query's grid (grid-stride over the rays) */
#define DMT_NOALLOC_SPILL_DEPTH  48u    /* spilled stack entries per thread */
dmt_status dmt_trace_closest_noalloc(dmt_scene* scene, const float* d_rays, uint32_t n, dmt_hit* d_out);

/* Counters of the frames rendered since the last reset. */
dmt_status dmt_scene_stats(dmt_scene* scene, dmt_stats* out);
dmt_status dmt_scene_reset_stats(dmt_scene* scene);
/* Re-renders one frame with the counting traversal (outside any timed region)
 * and fills node_visits / tri_tests / sphere_tests. */
dmt_status dmt_scene_measure_traffic(dmt_scene* scene, const dmt_frame_params* frame, dmt_stats* out);
/* Per-event timing of the closest-hit and frame-completion kernel launches
 * on/off (default off). */
dmt_status dmt_scene_set_profiling(dmt_scene* scene, int32_t enabled);
/* Frame schedule (no counterpart in the reference, whose frame is one
 * dispatch): the first `wavefront_iterations` bounces of a frame run as
 * per-bounce kernels over the live path queue, then one frame-completion
 * launch runs every remaining path, and the pixels' later samples, to the end
 * with lanes refilled from the queue.  DMT_SCHEDULE_WAVEFRONT keeps per-bounce
 * kernels to the end.  Images are identical under every schedule. */
#define DMT_SCHEDULE_WAVEFRONT 0xFFFFFFFFu
dmt_status dmt_scene_set_schedule(dmt_scene* scene, uint32_t wavefront_iterations);
dmt_status dmt_scene_get_schedule(dmt_scene* scene, uint32_t* wavefront_iterations);

/* Moving meshes: new positions for the scene's mesh (the same vertex count and
 * triangles) with the triangle BVH refit on the device -- the triangle records
 * re-formed and every node's boxes recomputed bottom-up from the new vertices,
 * the tree's topology kept.  Replaces the host rebuild the reference engine
 * runs when a mesh changes (RebuildMeshResources, Engine.cpp:3752-3900;
 * CreateBLAS/CreateTLAS, SoftwareDevice.cpp:337-421) where the triangles
 * stay the same.  Images are those of a scene built from the new positions
 * (the closest hit does not depend on the tree); traversal cost grows as the
 * vertices move away from where the tree was built, and dmt_scene_create
 * restores it.  Waits for the scene's frames in flight. */
dmt_status dmt_scene_update_mesh(dmt_scene* scene, const float* positions, uint32_t vertex_count);

/* ---- Material parameters and the adjoint -----------------------------------
 * Per analytic primitive, 7 floats: albedo rgb (v1.xyz), roughness (v1.w),
 * emission rgb (v4.xyz) of the primitive wire format.  Replaces re-uploading
 * the prim buffer after a material edit (Engine.cpp:4743-5002 re-packs all
 * prims); geometry and acceleration data are unchanged. */
#define DMT_MATERIAL_FLOATS 7u
dmt_status dmt_scene_get_materials(dmt_scene* scene, float* out, size_t float_count);
/* Waits for the scene's frames in flight, then updates the device prims. */
dmt_status dmt_scene_set_materials(dmt_scene* scene, const float* materials, size_t float_count);

/* No counterpart in the reference (it has no differentiable renderer).
 * Gradient of an image loss with respect to every analytic primitive's
 * material parameters, for the frame `frame` renders with reset_accum set
 * (image I = that frame's value): d_image_grad is a device array of
 * width*height float4 holding dLoss/dI (rgb; w ignored), d_param_grad a
 * device array of prim_count*7 floats that receives dLoss/dparams
 * (overwritten).  Path replay: the samples are re-traced from their seeds,
 * no tape is stored.  Albedo and emission gradients are exact derivatives of
 * the estimator with its random numbers fixed; the roughness gradient is the
 * derivative of the estimator away from visibility boundaries (decisions
 * held fixed).  With Russian roulette on, its probability is treated as a
 * constant (set DMT_FRAME_NO_RUSSIAN_ROULETTE for a gradient that finite
 * differences reproduce).  Limits: max_bounces <= 8, refract_bounces <= 8.
 * The camera-sample features replay as rendered (a spectral sample's
 * gradient carries its wavelength's weight); SDF clusters are traced and
 * shaded as rendered and carry no parameters (their hit distance and normal
 * are held fixed along the roughness tangent).  Frames with sky lights, PBR
 * maps or water are refused (DMT_ERR_UNSUPPORTED).  Stream-ordered on the device stream; d_image_grad must be complete when
 * the call is made. */
dmt_status dmt_adjoint(dmt_scene* scene, const dmt_frame_params* frame, const void* d_image_grad,
                       void* d_param_grad);

/* ---- Batched persistent form ---------------------------------------------
 * No one-for-one counterpart in the reference: its engine renders one scene
 * per frame, one PushConstants + Dispatch("pathtrace") each
 * (Engine.cpp:9577-9582 -> SoftwareDevice.cpp:105-136 -> RunPathTraceKernel).
 * A batch renders a frame of many scenes, each with its own camera and
 * resolution, in one persistent-kernel launch fed by a device work queue of
 * 64-pixel tiles; frames replay a captured hipGraph.  Scene i's image after
 * dmt_batch_render(b, 0, n) equals dmt_render_frames(scene_i, &frames[i], n).
 */
typedef struct dmt_batch dmt_batch;

typedef struct dmt_batch_params {
    uint32_t tile_w;            /* tile width in pixels, tiles hold 64 pixels:
                                   8, 16, 32 or 64 (0 = 8)                    */
    uint32_t tiles_per_fetch;   /* tiles a wave takes per queue atomic, 1..16 (0 = 1) */
    uint32_t blocks_per_cu;     /* persistent workgroups of 256 threads per CU,
                                   1..5 (0 = 4); stack_lds 16 holds 5 per CU,
                                   24 holds 4, 32 holds 3                     */
    uint32_t stack_lds;         /* LDS traversal-stack entries per thread:
                                   16, 24 or 32 (0 = 16)                      */
    uint32_t plain_launches;    /* 0 = replay a captured hipGraph per frame
                                   (default), 1 = launch the kernels directly */
    uint32_t reserved[3];
} dmt_batch_params;

/* Groups count scenes (distinct, on one device) with their frame settings.
 * Each scene's accumulation buffer is (re)sized to its frame's resolution;
 * while the batch lives the scene cannot be rendered at another resolution,
 * and dmt_scene_destroy of a member is deferred to dmt_batch_destroy.
 * params may be null (defaults).  Frames may use the camera-sample features
 * (each scene its own). */
dmt_status dmt_batch_create(dmt_device* dev, dmt_scene* const* scenes, const dmt_frame_params* frames,
                            uint32_t count, const dmt_batch_params* params, dmt_batch** out_batch);
void       dmt_batch_destroy(dmt_batch* batch);
/* Renders frames first_frame .. first_frame + frame_count - 1 of every scene:
 * frame k of scene i uses frame_index = frames[i].frame_index + k and
 * resets the accumulation at k == 0 when frames[i].reset_accum is set.
 * Stream-ordered after earlier work on the member scenes. */
dmt_status dmt_batch_render(dmt_batch* batch, uint32_t first_frame, uint32_t frame_count);
dmt_status dmt_batch_set_params(dmt_batch* batch, const dmt_batch_params* params);
dmt_status dmt_batch_get_params(const dmt_batch* batch, dmt_batch_params* out);
/* Times candidate parameters (coordinate descent over blocks_per_cu,
 * stack_lds, tile_w, tiles_per_fetch; frames_per_trial timed frames each,
 * rendered into scratch buffers so the scenes' images are untouched),
 * keeps the fastest and returns it with its time per frame. */
dmt_status dmt_batch_autotune(dmt_batch* batch, uint32_t frames_per_trial, dmt_batch_params* out_best,
                              float* out_ms_per_frame);
/* Ray counts of the batch frames rendered since the last reset. */
dmt_status dmt_batch_stats(dmt_batch* batch, dmt_stats* out);
dmt_status dmt_batch_reset_stats(dmt_batch* batch);

/* Last error message of the calling thread ("" if none). */
const char* dmt_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* DMT_H */
