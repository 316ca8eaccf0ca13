/* CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement of the reference path tracer (shaders/PathTrace.slang
 * main() with the off-path features disabled, plus the host builders in
 * src/renderer/) used as the correctness checker of the HIP path.  Nothing in
 * the product links, loads or calls this library; only tests/, the smoke
 * entry and bench.py's CPU-baseline leg do.
 *
 * Data layouts are the reference wire formats:
 *   prims   7 float4 (112 B) per analytic primitive (src/engine/Engine.cpp:4924-4994)
 *   lights  5 float4 (80 B) per analytic light      (src/engine/Engine.cpp:5235-5300)
 *   mesh    float[3V] positions, uint32[3T] indices (shaders/PathTrace.slang:44-45)
 *   env     float[3*W*H] linear RGB, row-major, v = 0 at the top (theta = 0)
 */
#ifndef DMT_ORACLE_H
#define DMT_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene_desc {
    const float*    prims;      uint32_t prim_count;
    const float*    lights;     uint32_t light_count;
    const float*    positions;  uint32_t vertex_count;
    const uint32_t* indices;    uint32_t index_count;
    const float*    env_rgb;    uint32_t env_width;  uint32_t env_height;
    float           env_intensity;
    uint32_t        reserved;
    /* optional importance CDF of the env map (as the engine uploads it;
     * NULL: built from env_rgb) */
    const float*    env_marginal;
    const float*    env_conditional;
    double          env_total_luminance;
    /* optional moon texture (RGBA floats) of the moon disc */
    const float*    moon_rgba;  uint32_t moon_width; uint32_t moon_height;
    /* optional SDF clusters, 112 floats each (the engine's sdf_clusters
     * records, Engine.cpp:5076-5190), traced after the analytic prims */
    const float*    sdf_clusters; uint32_t sdf_cluster_count;
    /* optional PBR material atlas as the engine builds it (RGBA8, 384 wide:
     * a 256-square base image per tile band and its mips in the right
     * column, 256 rows per tile; PathTrace.slang:1435-1605), sampled as
     * byte * (1/255); and per-vertex mesh UVs (2 floats each) */
    const uint8_t*  pbr_atlas;  uint32_t pbr_atlas_width; uint32_t pbr_atlas_height;
    const float*    mesh_uvs;
} orc_scene_desc;

typedef struct orc_frame {
    float    pos_fovtan[4];
    float    fwd_aspect[4];
    float    right_xyz[4];
    float    up_xyz[4];
    uint32_t width, height;
    uint32_t frame_index;
    uint32_t spp;
    uint32_t max_bounces;
    uint32_t reset_accum;
    uint32_t mis_enabled;
    uint32_t light_tree_enabled;
    uint32_t refract_bounces;
    float    firefly_clamp;
    float    ema_alpha;
    uint32_t flags;      /* bit 0: no Russian roulette (same bits as dmt_frame_params) */
    /* Camera-sample features (same fields as dmt_frame_params):
     * thin lens (PathTrace.slang:5779-5816), shutter time (:5761-5773),
     * hero wavelength (:5737-5759), mesh translation (:2461-2487). */
    float    dof[4];     /* aperture radius (0 = pinhole), focal distance, blades, blade rotation (rad) */
    float    cat_eye;    /* dof_bokeh_params.x */
    uint32_t motion_blur;
    uint32_t spectral;
    float    cauchy_c;   /* spectral_params.y, um^2 */
    float    mesh_shift_prev[3];
    uint32_t mesh_shift_active;
    float    mesh_shift_curr[3];
    uint32_t reserved;
    /* Fixed jitter of sample 0 (flags bit 1) and the sky lights (same fields
     * as dmt_frame_params): halton_jitter (PathTrace.slang:5730), the sun and
     * moon of the gradient sky (:3112, :3212, :6546-6667), the HDRI cluster
     * lights of an env map (:3679-3768, :6493-6518). */
    float    jitter[2];
    float    sun[4];
    float    sun_extra[4];
    float    sun_disc_brightness;
    float    moon[4];
    float    moon_extra[4];
    uint32_t hdri_light_count;
    float    hdri_dir[8][4];
    float    hdri_col[8][4];
    /* SDF sphere trace: step cap and surface distance (sdf_params.y,
     * sdf_params_f.x), read when the scene has SDF clusters */
    uint32_t sdf_max_iters;
    float    sdf_epsilon;
    /* atlas tiles of the mesh's albedo, normal, roughness and metallic maps
     * (mesh_tex_indices; 0xFFFFFFFF = none) */
    uint32_t mesh_tex[4];
    /* Water (water_params0, water_params1, water_state): Beer's-law absorption
     * rgb (1/m), the surface's index, the wave normal's scale (cycles/m),
     * amplitude, speed and time, the camera's in-water flag and the deep-water
     * tint (PathTrace.slang:5089-5296, 5847-5872, 6014-6050, 6768-6906) */
    float    water_absorb[3];
    float    water_ior;
    float    wave[4];
    uint32_t camera_in_water;
    float    deep_color[3];
} orc_frame;

typedef struct orc_hit {
    float    t;
    float    normal[3];
    uint32_t hit;        /* 0 miss, 1 analytic, 2 mesh */
    uint32_t prim;       /* analytic: index in input order; mesh: triangle index */
    uint32_t mat;
    uint32_t user_id;    /* prims[v2.w] for analytic, 0 for mesh (PathTrace.slang:2659) */
} orc_hit;

typedef struct orc_counts {
    uint64_t primary;    /* b == 0 extension queries */
    uint64_t extension;  /* b >= 1 extension queries */
    uint64_t shadow;     /* every traceScene step inside transmittance() */
    uint64_t node_visits;
    uint64_t prim_tests;
} orc_counts;

typedef struct orc_scene orc_scene;

/* Build the oracle's own acceleration data (median BVHs, light tree, env CDF). */
int  orc_scene_create(const orc_scene_desc* desc, orc_scene** out);
void orc_scene_destroy(orc_scene* s);

/* Render one frame: per pixel in the selected rows/columns, run the spp loop,
 * average, sanitize and fold into accum (float4 per pixel, running mean).
 * precision: 32 = float arithmetic (parity oracle), 64 = double (convergence).
 * Rows y = row_begin, row_begin + row_step, ... < row_end; columns x = 0,
 * col_step, ...  threads <= 0 means hardware_concurrency. */
int orc_render(const orc_scene* s, const orc_frame* f, int precision, int threads,
               uint32_t row_begin, uint32_t row_end, uint32_t row_step, uint32_t col_step,
               float* accum_rgba, orc_counts* counts);

/* fp64 tangent-linear oracle: renders the frame in dual numbers along the
 * direction of material parameter `param` (0-2 albedo rgb, 3 roughness, 4-6
 * emission rgb) of analytic prim `prim` and returns
 * sum_p sum_c dimage[p].c * dI[p].c/dparam in *out_dloss, where I is the frame
 * value (spp mean) and dimage is width*height float4 (rgb used).  Decisions
 * follow the fp64 values, so this is the exact derivative of the fp64
 * estimator with its random numbers fixed.  out_dimage (optional, float4 per
 * pixel) receives dI/dparam. */
int orc_tangent(orc_scene* s, const orc_frame* f, uint32_t prim, uint32_t param, const float* dimage,
                int threads, double* out_dloss, float* out_dimage);
/* The fp64 frame value (spp mean, sanitised) of every pixel, rgb doubles. */
int orc_render_f64_image(orc_scene* s, const orc_frame* f, int threads, double* out_rgb);

/* skyColor of a missed ray for n unit directions (3 floats each) under the
 * frame's sky lights: the env map, or the gradient sky with the sun and moon
 * discs; out: 3 floats each. */
int orc_sky(const orc_scene* s, const orc_frame* f, const float* dirs, uint32_t n, float* out_rgb);
/* sampleSunDisc (PathTraceMath.slang:55-67) around dir with n successive
 * seeds seed0, seed0 + 1, ...: out 3 floats each. */
void orc_sample_sun_disc(const float* dir, uint32_t seed0, uint32_t n, float* out);

/* Closest-hit query (traceScene) for n rays: rays are 8 floats each
 * (ox, oy, oz, unused, dx, dy, dz, unused). */
int orc_trace(const orc_scene* s, const float* rays, uint32_t n, orc_hit* out);
/* Signed distance and unit normal (sdfClusterDist, normalize(sdfClusterNormal),
 * SdfPrimitives.slang:377-491) of one 112-float SDF cluster at n points
 * (3 floats each): out 4 floats each (d, nx, ny, nz). */
void orc_sdf_eval(const float* cluster, const float* points, uint32_t n, float* out);
/* sdfClusterTrace (SdfPrimitives.slang:1330-1444) of one cluster for n rays
 * (8 floats each) with t_min 1e-3 and t_max 1e30: out 4 floats each
 * (t or -1 for none, normal xyz). */
void orc_sdf_trace(const float* cluster, const float* rays, uint32_t n, uint32_t max_iters, float epsilon,
                   float* out);
/* Same through an exhaustive linear scan (no BVH). */
int orc_trace_bruteforce(const orc_scene* s, const float* rays, uint32_t n, orc_hit* out);
/* traceScene at shutter time `time` (spheres at lerp(v3, v0, time)); with
 * shift_prev and shift_curr non-null the mesh is translated by
 * lerp(shift_prev, shift_curr, time).  brute != 0: exhaustive scan. */
int orc_trace_at(const orc_scene* s, const float* rays, uint32_t n, float time, const float* shift_prev,
                 const float* shift_curr, int brute, orc_hit* out);

/* Denoiser G-buffer of a frame (PathTrace.slang:5463-5648): the pixel-centre
 * camera ray of every pixel traced once; 16 floats per pixel: depth, motion.xy,
 * normal.xyz, albedo.xyz, specular albedo (F0).xyz, roughness, specular hit
 * distance, 0, 0.  vp_curr / vp_prev are column-major (clip = M * (p, 1)). */
int orc_gbuffer(const orc_scene* s, const orc_frame* f, const float* vp_curr, const float* vp_prev, float* out);

/* PBR atlas probes (fp32): samplePbrTileLinear of `tile` at (u, v, lod)
 * triples -> 4 floats each; the hit's PBR surface for n rays (8 floats each)
 * with the given ray-cone width -> 16 floats each: u, v, uv_scale, lod,
 * textured albedo rgb, roughness, material, shading normal xyz, t, kind
 * (0 miss), 0, 0; applyNormalMap of n normals by n tangent-space vectors;
 * srgbToLinear of n values. */
int  orc_pbr_sample(const orc_scene* s, uint32_t tile, const float* uvl, uint32_t n, float* out);
int  orc_pbr_surface(const orc_scene* s, const orc_frame* f, const float* rays, uint32_t n, float cone_width,
                     float* out);
void orc_pbr_normal_map(const float* normals, const float* ts, uint32_t n, float* out);
void orc_pbr_srgb(const float* c, uint32_t n, float* out);

/* cloud_value_noise (PathTraceCloud.slang:31-69) at n points (3 floats
 * each), and waterShadingNormal (PathTrace.slang:5231-5270) of the frame's
 * wave parameters at n points with macro normals (3 floats each). */
void orc_value_noise(const float* p, uint32_t n, float* out);
int  orc_water_normal(const orc_frame* f, const float* p, const float* n_macro, uint32_t n, float* out);

/* Host builders, exposed for tests. */
double orc_build_env_cdf(const float* rgb, uint32_t W, uint32_t H,
                         float* out_marginal, float* out_conditional);
/* nodes: 16 floats (64 B) per node, at most 2*light_count - 1 nodes. */
uint32_t orc_build_light_tree(const float* lights, uint32_t light_count, float* out_nodes);
/* Median BVH over triangles (TriangleBvh): nodes 8 x u32 each, perm[T].  Returns node count. */
uint32_t orc_build_tri_bvh(const float* positions, uint32_t vertex_count,
                           const uint32_t* indices, uint32_t index_count,
                           uint32_t* out_nodes, uint32_t* out_perm);

/* Probes of the shading helpers (fp32), for the pinning tests.
 * env sample: xi 2 floats each -> 8 floats (dir.xyz, colour.rgb, pdf, 0). */
int orc_env_sample(const orc_scene* s, const float* xi, uint32_t n, float* out);
int orc_env_pdf(const orc_scene* s, const float* dirs, uint32_t n, float* out);
/* Exhaustive light-tree pick PMF at point p with facing normal nf. */
int orc_light_pmf(const orc_scene* s, const float* p, const float* nf, float* pmf);

/* Math kernels exposed for the pinning tests (fp32 Cephes restatement).
 * fn: 0 sin, 1 cos, 2 acos, 3 atan2(a, b), 4 exp,
 *     5/6/7 spectral weight r/g/b at wavelength a (nm),
 *     8 Cauchy index at wavelength a (nm) of base index b (C = 0.0042 um^2). */
void orc_math_f32(int fn, const float* a, const float* b, float* out, uint32_t n);

/* Lens samples (fp32): u 2 floats each -> 2 floats.  blades < 3: the
 * concentric disk (ptConcentricDisk); else the area-uniform polygon
 * (ptPolygonSample) rotated by rot.  cat_eye > 0 then applies ptCatEye with
 * the screen position uv (2 floats per sample; may be null when cat_eye is 0). */
int orc_lens_sample(const float* u, const float* uv, uint32_t n, int blades, float rot, float cat_eye, float* out);

#ifdef __cplusplus
}
#endif
#endif
