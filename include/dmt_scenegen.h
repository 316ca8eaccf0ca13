/*
 * dmt_scenegen.h -- seeded synthetic workloads for the path tracer.
 *
 * No counterpart in the reference (its scenes come from console commands and
 * glTF); these generators produce the benchmark and parity workloads in the
 * reference wire formats of dmt.h.  Host only, no device needed.
 *
 *   config 0  64 spheres (centres U[-4,4]^3, radii U[0.2,1]) above the plane
 *             y = -5; materials by seeded shuffle of i mod 3 (Lambert albedo
 *             U[0.1,0.9]^3 / metal roughness U[0,0.5] / dielectric IOR 1.5);
 *             one quad light (emission 10); all-zero env (black sky);
 *             256x256, 1 spp, 4 bounces.
 *   config 1  icosphere (radius 3) subdivided 8x, displaced along the normal by
 *             seeded value noise (amplitude 0.1), plus 200k random triangles
 *             (edges <= 0.05) uniform in [-5,5]^3; one quad light; 512x256
 *             env with lognormal texels exp(N(0,1)); 1920x1080, 1 spp, 8 bounces.
 *   config 2  4,096 analytic objects over the plane y = -5: 3,584 spheres
 *             (centres U[-4.5,4.5]^3, radii U[0.05,0.3], config 0 materials)
 *             and 512 SDF clusters (boxes, rounded boxes, a box united with a
 *             sphere, a sphere smoothly united with a capsule, a box with a
 *             sphere carved out, tori; sizes ~0.1-0.3, centres U[-4.5,4.5]^3,
 *             config 0 materials), a 500k-triangle mesh
 *             (icosphere-7 of radius 2.5 + 172,320 random small triangles in
 *             [-5,5]^3), lit only by 4,096 emissive triangles (light type 4,
 *             edges ~0.3, facing down, centres in [-6,6] x [5,7] x [-6,6],
 *             emission exp(N(0,1)) per channel); all-zero env; 1280x720,
 *             4 spp, 6 bounces.
 *   config 4  config 0 with 256 spheres, 256 distinct materials (each
 *             sphere also emits U[0.02, 0.3] per channel); 512x512, 16 spp,
 *             4 bounces.
 *   config 1 with options.spheres > 0 adds config 0's spheres and plane (used
 *             by the batched workload, see dmt_batch_workload_scene).
 * Camera for all: at (0, 2, 12) looking at the origin, vertical fov 50 deg.
 */
#ifndef DMT_SCENEGEN_H
#define DMT_SCENEGEN_H
#include "dmt.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct dmt_gen_options {
    int32_t  subdivisions;   /* icosphere subdivisions (< 0: config default)   */
    int32_t  random_tris;    /* random small triangles (< 0: config default)   */
    int32_t  spheres;        /* analytic spheres (< 0: config default)         */
    uint32_t width, height;  /* 0: config default                              */
    uint32_t env_width, env_height; /* 0: config default                       */
    int32_t  lights;         /* emissive triangle lights (config 2; < 0: default) */
    int32_t  sdf_clusters;   /* SDF clusters (config 2; < 0: default 512)      */
} dmt_gen_options;

typedef struct dmt_generated_scene {
    float*           prims;      uint32_t prim_count;
    float*           lights;     uint32_t light_count;
    float*           positions;  uint32_t vertex_count;
    uint32_t*        indices;    uint32_t index_count;
    float*           env_rgb;    uint32_t env_width;  uint32_t env_height;
    float            env_intensity;
    dmt_frame_params frame;      /* the config's camera and render settings  */
    float*           sdf_clusters; uint32_t sdf_cluster_count;   /* DMT_SDF_CLUSTER_FLOATS each */
} dmt_generated_scene;

dmt_status dmt_scene_generate(uint32_t config, uint64_t seed, const dmt_gen_options* options,
                              dmt_generated_scene** out);
void       dmt_scene_generated_free(dmt_generated_scene* scene);

/* Scene `index` of the batched workload: resolution drawn uniformly from
 * {128, 256, 512, 1024} squared; triangle count log-uniform in [10k, 2M],
 * split into the finest icosphere that fits plus random triangles of the
 * config 1 generator; odd indices also get config 0's 64 spheres and ground
 * plane.  Fills the options and seed for dmt_scene_generate(1, ...); the
 * frame is that scene's camera with 1 spp and *out_max_bounces (6). */
dmt_status dmt_batch_workload_scene(uint64_t seed, uint32_t index, dmt_gen_options* out_options,
                                    uint64_t* out_scene_seed, uint32_t* out_max_bounces);

/* Camera basis of the reference Camera (src/renderer/Camera.h) looking from
 * eye to target; fills pos_fovtan / fwd_aspect / right_xyz / up_xyz. */
dmt_status dmt_camera_look_at(const float eye[3], const float target[3], float fov_y_deg,
                              uint32_t width, uint32_t height, dmt_frame_params* frame);

/* Host builders behind dmt_scene_create, exposed so their output can be
 * checked without a device.
 *
 * Binned-SAH BVH over triangles, collapsed to 4-wide 128 B nodes (replaces
 * TriangleBvh::Build, src/renderer/TriangleBvh.cpp:33-91).  Two-call pattern:
 * pass null out_nodes/out_order to get the counts.  out_nodes receives
 * node_count * 128 B, out_order the leaf-order triangle ids. */
dmt_status dmt_build_bvh4(const float* positions, uint32_t vertex_count, const uint32_t* indices,
                          uint32_t index_count, uint32_t sah_bins, uint32_t max_leaf, void* out_nodes,
                          void* out_order, uint32_t* out_node_count, uint32_t* out_depth);
/* BuildLightTree (src/renderer/LightTree.h:122-123): 16 floats per node,
 * at most 2 * light_count - 1 nodes; returns the node count. */
uint32_t dmt_build_light_tree(const float* lights, uint32_t light_count, float* out_nodes);
/* BuildEnvCdf (src/renderer/HdrImage.h:54-58): marginal[H], conditional[W*H];
 * returns the sin-weighted luminance total. */
double dmt_build_env_cdf(const float* rgb, uint32_t width, uint32_t height, float* out_marginal,
                         float* out_conditional);
/* The same with the pixels mask[] != 0 left out of the luminance (the
 * engine's light_mask of extracted HDRI clusters, Engine.cpp:4559-4563). */
double dmt_build_env_cdf_masked(const float* rgb, uint32_t width, uint32_t height, const uint8_t* mask,
                                float* out_marginal, float* out_conditional);
/* The engine's HDRI light extraction at env-map load (ReloadEnvMap,
 * Engine.cpp:4369-4532): pixels at or above the top_frac luminance quantile
 * (r_hdri_extract_percentile, 0.005 by default; floored at 5 % of the peak
 * and at 0.5, none when the peak is <= 0.5), 4-connected with wrap-around
 * columns, the 8 brightest components of >= 3 pixels as directional lights:
 * dir[i] = (unit direction, pick probability), col[i] = (rgb flux, 0), as
 * dmt_frame_params.hdri_dir / hdri_col take them; out_mask (width*height)
 * marks their pixels for dmt_build_env_cdf_masked.  Returns the number of
 * lights. */
uint32_t dmt_env_extract_lights(const float* rgb, uint32_t width, uint32_t height, double top_frac,
                                float dir[8][4], float col[8][4], uint8_t* out_mask);

#ifdef __cplusplus
}
#endif
#endif /* DMT_SCENEGEN_H */
