// Host C++ under AddressSanitizer and UndefinedBehaviorSanitizer.
//
// Drives the code that manages memory by hand on the host -- the scene
// builders and validation (builders.cpp, prep.cpp), the workload generators
// (scenegen.cpp), the RHI adapter's push decoding and host resources
// (hip_device.cpp, without a GPU) and the CPU oracle (pt_oracle.cpp) --
// through their entry points with ordinary and edge-case inputs.  Built and
// run by tests/test_host_sanitizers.py; any sanitizer report aborts the run.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "dmt.h"
#include "dmt_rhi.h"
#include "dmt_scenegen.h"
#include "oracle.h"
#include "prep.h"

static_assert(sizeof(orc_frame) == sizeof(dmt_frame_params), "the oracle frame mirrors dmt_frame_params");

namespace {

int g_failures = 0;

void check(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "FAILED: %s\n", what);
        ++g_failures;
    }
}

orc_frame to_orc(const dmt_frame_params& f) {
    orc_frame o;
    std::memcpy(&o, &f, sizeof(o));
    return o;
}

dmt_generated_scene* generate(uint32_t config, uint64_t seed, dmt_gen_options opt) {
    dmt_generated_scene* g = nullptr;
    check(dmt_scene_generate(config, seed, &opt, &g) == DMT_OK && g, "dmt_scene_generate");
    return g;
}

dmt_gen_options small_options(uint32_t w, uint32_t h) {
    dmt_gen_options o;
    o.subdivisions = 2;
    o.random_tris = 300;
    o.spheres = -1;
    o.width = w;
    o.height = h;
    o.env_width = 32;
    o.env_height = 16;
    o.lights = 40;
    o.sdf_clusters = 24;
    return o;
}

// The host builders and the scene validation on one generated scene.
void builders(const dmt_generated_scene& g) {
    uint32_t nodes = 0, depth = 0;
    check(dmt_build_bvh4(g.positions, g.vertex_count, g.indices, g.index_count, 16, 4, nullptr, nullptr, &nodes,
                         &depth) == DMT_OK, "bvh4 count");
    std::vector<unsigned char> nbuf((size_t)nodes * 128u + 1u);
    std::vector<uint32_t> order(g.index_count / 3u + 1u);
    check(dmt_build_bvh4(g.positions, g.vertex_count, g.indices, g.index_count, 16, 4, nbuf.data(), order.data(),
                         &nodes, &depth) == DMT_OK, "bvh4 build");
    if (g.light_count) {
        std::vector<float> t((2u * g.light_count) * 16u);
        check(dmt_build_light_tree(g.lights, g.light_count, t.data()) == 2u * g.light_count - 1u, "light tree");
    }
    if (g.env_width) {
        std::vector<float> m(g.env_height), c((size_t)g.env_width * g.env_height);
        dmt_build_env_cdf(g.env_rgb, g.env_width, g.env_height, m.data(), c.data());
    }
    dmt::PreparedScene S;
    const std::string err = dmt::prepare_scene(g.prims, g.prim_count, g.lights, g.light_count, g.positions,
                                               g.vertex_count, g.indices, g.index_count, g.env_rgb, g.env_width,
                                               g.env_height, g.env_intensity, 16, 4, S);
    check(err.empty(), "prepare_scene of a generated scene");
    if (g.sdf_cluster_count) {
        bool unsupported = false;
        check(dmt::prepare_sdf(g.sdf_clusters, g.sdf_cluster_count, S, &unsupported).empty(), "prepare_sdf");
    }
}

// Edge cases of the builders and of validation: empty, single, duplicated
// and degenerate inputs, out-of-range indices, non-finite values.
void builder_edges() {
    const float tri[9] = {0, 0, 0, 1, 0, 0, 0, 1, 0};
    const uint32_t idx[3] = {0, 1, 2}, bad[3] = {0, 1, 3};
    uint32_t nodes = 0, depth = 0;
    check(dmt_build_bvh4(nullptr, 0, nullptr, 0, 16, 4, nullptr, nullptr, &nodes, &depth) == DMT_OK, "empty bvh4");
    check(dmt_build_bvh4(tri, 3, idx, 3, 16, 4, nullptr, nullptr, &nodes, &depth) == DMT_OK, "one-tri bvh4");
    check(dmt_build_bvh4(tri, 3, bad, 3, 16, 4, nullptr, nullptr, &nodes, &depth) == DMT_ERR_INVALID_ARG,
          "bvh4 index out of range");
    check(dmt_build_bvh4(tri, 3, idx, 2, 16, 4, nullptr, nullptr, &nodes, &depth) == DMT_ERR_INVALID_ARG,
          "bvh4 ragged index count");
    // 2,000 copies of one triangle: every split degenerates
    std::vector<uint32_t> dup(6000);
    for (size_t i = 0; i < dup.size(); ++i) dup[i] = (uint32_t)(i % 3u);
    check(dmt_build_bvh4(tri, 3, dup.data(), (uint32_t)dup.size(), 16, 4, nullptr, nullptr, &nodes, &depth) == DMT_OK,
          "duplicated-triangle bvh4 count");
    std::vector<unsigned char> nb((size_t)nodes * 128u);
    std::vector<uint32_t> ord(2000);
    check(dmt_build_bvh4(tri, 3, dup.data(), (uint32_t)dup.size(), 16, 4, nb.data(), ord.data(), &nodes, &depth) ==
              DMT_OK, "duplicated-triangle bvh4");
    // light trees of 0, 1 and coincident lights of every type
    float node[16 * 64];
    check(dmt_build_light_tree(nullptr, 0, node) == 0u, "no lights");
    std::vector<float> lights(20u * 32u, 0.0f);
    for (uint32_t i = 0; i < 32; ++i) {
        float* l = &lights[20u * i];
        l[3] = 0.1f;
        l[4] = l[5] = l[6] = 1.0f;
        l[7] = (float)(i % 5u);
        l[9] = -1.0f;
        l[11] = 0.5f;
        l[12] = 0.2f;
        l[15] = 0.3f;
        l[19] = 1.0f;
    }
    for (uint32_t n : {1u, 2u, 3u, 32u}) check(dmt_build_light_tree(lights.data(), n, node) == 2u * n - 1u, "light tree n");
    // env CDFs of a single texel, an all-zero map and a masked map
    const float one[3] = {1, 2, 3};
    float m1[1], c1[1];
    dmt_build_env_cdf(one, 1, 1, m1, c1);
    std::vector<float> zero(3u * 8u * 4u, 0.0f), m(4), c(32);
    check(dmt_build_env_cdf(zero.data(), 8, 4, m.data(), c.data()) == 0.0, "zero env total");
    std::vector<float> env(3u * 64u * 32u);
    std::mt19937 rng(7);
    std::lognormal_distribution<float> ln(0.0f, 1.0f);
    for (float& v : env) v = ln(rng);
    for (int k = 0; k < 3; ++k) env[3u * (5u * 64u + 63u) + k] = env[3u * (5u * 64u) + k] = 500.0f;   // wraps around
    float dir[8][4], col[8][4];
    std::vector<uint8_t> mask(64u * 32u);
    const uint32_t nl = dmt_env_extract_lights(env.data(), 64, 32, 0.005, dir, col, mask.data());
    std::vector<float> mm(32), cc(64u * 32u);
    dmt_build_env_cdf_masked(env.data(), 64, 32, mask.data(), mm.data(), cc.data());
    check(nl <= 8u, "extracted lights");
    // validation refuses bad scenes
    dmt::PreparedScene S;
    float prim[28] = {0};
    prim[3] = 1.0f;
    prim[4] = prim[5] = prim[6] = 0.5f;
    prim[8] = 0.0f;
    prim[9] = 7.0f;   // material 7: invalid
    prim[23] = 1.0f;
    check(!dmt::prepare_scene(prim, 1, nullptr, 0, nullptr, 0, nullptr, 0, nullptr, 0, 0, 1.0f, 16, 4, S).empty(),
          "bad material refused");
    prim[9] = 0.0f;
    prim[0] = NAN;
    check(!dmt::prepare_scene(prim, 1, nullptr, 0, nullptr, 0, nullptr, 0, nullptr, 0, 0, 1.0f, 16, 4, S).empty(),
          "NaN prim refused");
    check(!dmt::prepare_scene(nullptr, 0, nullptr, 0, tri, 3, bad, 3, nullptr, 0, 0, 1.0f, 16, 4, S).empty(),
          "bad mesh index refused");
    // SDF records: an unknown shape and a child index past the node count
    std::vector<float> sdf(112u, 0.0f);
    sdf[0] = -1; sdf[1] = -1; sdf[2] = -1; sdf[4] = 1; sdf[5] = 1; sdf[6] = 1;
    sdf[3] = 1.0f;                        // one node
    sdf[16] = 0.0f; sdf[17] = 99.0f;      // leaf with shape 99
    bool unsupported = false;
    dmt::PreparedScene S2;
    check(!dmt::prepare_sdf(sdf.data(), 1, S2, &unsupported).empty() || unsupported, "bad SDF refused");
}

// PtPush decoding at every size, on random bytes, and the device's host
// resources without a GPU.
void rhi() {
    std::vector<unsigned char> push(2048);
    std::mt19937 rng(11);
    for (int trial = 0; trial < 200; ++trial) {
        for (auto& b : push) b = (unsigned char)rng();
        const size_t size = (trial * 37u) % (push.size() + 1u);
        dmt::rhi::PathtracePush P{};
        std::string err;
        dmt::rhi::DecodePathtracePush(push.data(), size, trial & 1, &P, &err);
    }
    for (size_t size : {0ul, 1ul, 111ul, 112ul, 1791ul, 1792ul, 2048ul}) {
        std::fill(push.begin(), push.end(), 0);
        dmt::rhi::PathtracePush P{};
        std::string err;
        dmt::rhi::DecodePathtracePush(push.data(), size, true, &P, &err);
    }
    using namespace dmt::rhi;
    HipDevice dev(HipDeviceOptions{});
    if (!dev.IsDeviceLost()) return;   // a GPU is present: the GPU tests cover the device
    const BufferHandle b = dev.CreateBuffer(BufferDesc{64, BufferUsage::Storage, "b"});
    check((bool)b, "host buffer");
    std::vector<float> data(8, 1.0f), back(16, 0.0f);
    const std::string before = dev.LastError();
    dev.WriteBuffer(b, data.data(), 32, 16);
    check(dev.LastError() == before, "write in range");
    dev.WriteBuffer(b, data.data(), 32, 40);
    check(dev.LastError().find("exceeds buffer size") != std::string::npos, "write out of range refused");
    check(dev.ReadbackBuffer(b, back.data(), 64) && back[4] == 1.0f && back[12] == 0.0f, "readback");
    check(!dev.ReadbackBuffer(b, back.data(), 65), "oversized readback refused");
    const float pos[9] = {0, 0, 0, 1, 0, 0, 0, 1, 0};
    const uint32_t idx[3] = {0, 1, 2}, bad[3] = {0, 1, 5};
    const AccelStructHandle blas = dev.CreateBLAS(BLASDesc{pos, 3, idx, 3, "m"});
    check((bool)blas, "blas");
    check(!dev.CreateBLAS(BLASDesc{pos, 3, bad, 3, "m"}), "bad blas refused");
    TLASInstance inst[2] = {};
    inst[0].blas = blas;
    inst[1].blas = AccelStructHandle{999999u};
    check((bool)dev.CreateTLAS(TLASDesc{std::span<const TLASInstance>(inst, 2), "t"}), "tlas skips unknown blas");
    check(!dev.CreateTexture(TextureDesc{4, 4, TextureFormat::RGBA32F, TextureUsage::Storage, "x"}),
          "textures need device memory");
    // a dispatch without a device is refused, not run
    HipCommandBuffer* cb = dev.AcquireCommandBuffer();
    const PipelineHandle p = dev.CreateComputePipeline(ComputePipelineDesc{"pathtrace", {}, "pt"});
    if (cb) {
        cb->BindComputePipeline(p);
        cb->BindBuffer(3, b);
        cb->BindAccelStruct(2, blas);
        std::fill(push.begin(), push.end(), 0);
        cb->PushConstants(push.data(), 1792);
        cb->Dispatch(1, 1, 1);
        dev.Submit(cb);
    }
    dev.DestroyAccelStruct(blas);
    dev.DestroyBuffer(b);
    dev.DestroyPipeline(p);
}

// The oracle: every render mode on small scenes of each generator.
void oracle(const dmt_generated_scene& g, bool tangent) {
    orc_scene_desc d{};
    d.prims = g.prims;
    d.prim_count = g.prim_count;
    d.lights = g.lights;
    d.light_count = g.light_count;
    d.positions = g.positions;
    d.vertex_count = g.vertex_count;
    d.indices = g.indices;
    d.index_count = g.index_count;
    d.env_rgb = g.env_rgb;
    d.env_width = g.env_width;
    d.env_height = g.env_height;
    d.env_intensity = g.env_intensity;
    d.sdf_clusters = g.sdf_clusters;
    d.sdf_cluster_count = g.sdf_cluster_count;
    orc_scene* s = nullptr;
    check(orc_scene_create(&d, &s) == 0 && s, "orc_scene_create");
    if (!s) return;
    orc_frame f = to_orc(g.frame);
    f.spp = 2;
    if (f.sdf_max_iters == 0) f.sdf_max_iters = 128, f.sdf_epsilon = 1e-4f;
    const uint32_t npx = f.width * f.height;
    std::vector<float> accum(4u * npx, 0.0f);
    orc_counts cnt{};
    check(orc_render(s, &f, 32, 4, 0, f.height, 1, 1, accum.data(), &cnt) == 0, "orc_render fp32");
    f.reset_accum = 0;
    f.frame_index++;
    f.ema_alpha = 0.5f;
    check(orc_render(s, &f, 64, 4, 0, f.height, 1, 1, accum.data(), &cnt) == 0, "orc_render fp64");
    // camera-sample features and the sky
    orc_frame x = f;
    x.dof[0] = 0.05f; x.dof[1] = 10.0f; x.dof[2] = 6.0f; x.cat_eye = 0.3f;
    x.motion_blur = 1u; x.spectral = 1u; x.cauchy_c = 0.0042f;
    x.flags |= DMT_FRAME_FIXED_JITTER;
    x.jitter[0] = 0.25f; x.jitter[1] = -0.125f;
    if (!g.env_width) {
        x.sun[1] = 0.8f; x.sun[2] = -0.6f;
        x.sun_extra[0] = 1.0f; x.sun_extra[1] = 1.0f; x.sun_extra[2] = 1.0f;
        x.sun_disc_brightness = 1.0f;
    }
    check(orc_render(s, &x, 32, 4, 0, f.height, 1, 1, accum.data(), &cnt) == 0, "orc_render features");
    std::vector<float> rays(8u * 64u);
    std::mt19937 rng(3);
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    for (uint32_t i = 0; i < 64; ++i) {
        float* r = &rays[8u * i];
        r[0] = 0.0f; r[1] = 2.0f; r[2] = 12.0f;
        const float dx = u(rng) * 0.4f, dy = u(rng) * 0.4f, dz = -1.0f;
        const float n = std::sqrt(dx * dx + dy * dy + dz * dz);
        r[4] = dx / n; r[5] = dy / n; r[6] = dz / n;
    }
    std::vector<orc_hit> h1(64), h2(64);
    orc_trace(s, rays.data(), 64, h1.data());
    orc_trace_bruteforce(s, rays.data(), 64, h2.data());
    for (uint32_t i = 0; i < 64; ++i) check(h1[i].hit == h2[i].hit && h1[i].t == h2[i].t, "bvh = brute force");
    std::vector<float> gb(16u * npx);
    float vp[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    orc_gbuffer(s, &f, vp, vp, gb.data());
    if (tangent) {
        std::vector<float> dimage(4u * npx, 1.0f);
        double dl = 0.0;
        for (uint32_t param : {0u, 3u, 6u}) orc_tangent(s, &f, 0, param, dimage.data(), 4, &dl, nullptr);
        std::vector<double> img(3u * npx);
        orc_render_f64_image(s, &f, 4, img.data());
    }
    orc_scene_destroy(s);
}

}  // namespace

int main() {
    builder_edges();
    rhi();
    for (uint32_t config : {0u, 1u, 2u, 4u}) {
        dmt_generated_scene* g = generate(config, 5u + config, small_options(40, 28));
        if (!g) continue;
        builders(*g);
        oracle(*g, config == 0u || config == 4u);
        dmt_scene_generated_free(g);
    }
    dmt_gen_options bo;
    uint64_t seed = 0;
    uint32_t bounces = 0;
    check(dmt_batch_workload_scene(1, 3, &bo, &seed, &bounces) == DMT_OK, "batch workload scene");
    std::printf("host sanitizers: %d failure(s)\n", g_failures);
    return g_failures ? 1 : 0;
}
