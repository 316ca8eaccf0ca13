// Seeded synthetic workloads (see dmt_scenegen.h for the configurations).
#include "../../../include/dmt_scenegen.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <unordered_map>
#include <vector>

namespace {

// splitmix64 stream: small, seedable, identical on every platform.
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double u01() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
    float uf(float a, float b) { return float(a + (b - a) * u01()); }
    double normal() {
        double u1 = u01(), u2 = u01();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

void put_prim(std::vector<float>& out, const float pos[3], float r, const float alb[3], float rough, uint32_t type,
              uint32_t mat, float ior, uint32_t user_id, const float em[3]) {
    float p[28] = {};
    p[0] = pos[0]; p[1] = pos[1]; p[2] = pos[2]; p[3] = r;
    p[4] = alb[0]; p[5] = alb[1]; p[6] = alb[2]; p[7] = rough;
    p[8] = float(type); p[9] = float(mat); p[10] = ior;
    std::memcpy(&p[11], &user_id, 4);
    p[12] = pos[0]; p[13] = pos[1]; p[14] = pos[2];
    p[16] = em[0]; p[17] = em[1]; p[18] = em[2];
    p[23] = 1.0f;   // identity orientation
    uint32_t none = 0xFFFFFFFFu;
    for (int k = 24; k < 28; ++k) std::memcpy(&p[k], &none, 4);
    out.insert(out.end(), p, p + 28);
}

void put_quad_light(std::vector<float>& out, const float pos[3], const float n[3], const float u[3], float v_half,
                    const float I[3]) {
    float L[20] = {};
    L[0] = pos[0]; L[1] = pos[1]; L[2] = pos[2]; L[3] = 0.0f;
    L[4] = I[0]; L[5] = I[1]; L[6] = I[2]; L[7] = 3.0f;
    L[8] = n[0]; L[9] = n[1]; L[10] = n[2]; L[11] = 0.0f;
    L[12] = u[0]; L[13] = u[1]; L[14] = u[2]; L[15] = v_half;
    L[19] = 1.0f;
    out.insert(out.end(), L, L + 20);
}

// Spheres over the ground plane, materials by a seeded shuffle of i mod 3;
// centres uniform in [-extent, extent]^3, radii uniform in [rmin, rmax].
void gen_spheres(Rng& rng, int n, std::vector<float>& prims, float extent = 4.0f, float rmin = 0.2f,
                 float rmax = 1.0f) {
    const float plane_n[3] = {0.0f, 1.0f, 0.0f}, grey[3] = {0.5f, 0.5f, 0.5f}, zero[3] = {0, 0, 0};
    put_prim(prims, plane_n, 5.0f, grey, 0.0f, 1u, 0u, 1.0f, 1u, zero);
    std::vector<uint32_t> mats((size_t)n);
    for (int i = 0; i < n; ++i) mats[(size_t)i] = (uint32_t)(i % 3);
    for (int i = n - 1; i > 0; --i) {
        int j = (int)(rng.next() % (uint64_t)(i + 1));
        std::swap(mats[(size_t)i], mats[(size_t)j]);
    }
    for (int i = 0; i < n; ++i) {
        float c[3] = {rng.uf(-extent, extent), rng.uf(-extent, extent), rng.uf(-extent, extent)};
        float r = rng.uf(rmin, rmax);
        float alb[3] = {rng.uf(0.1f, 0.9f), rng.uf(0.1f, 0.9f), rng.uf(0.1f, 0.9f)};
        float rough = rng.uf(0.0f, 0.5f);
        uint32_t m = mats[(size_t)i];
        if (m == 0u) rough = 0.0f;
        if (m == 2u) { alb[0] = alb[1] = alb[2] = 1.0f; rough = 0.0f; }
        put_prim(prims, c, r, alb, rough, 0u, m, m == 2u ? 1.5f : 1.0f, (uint32_t)(i + 2), zero);
    }
}

// An SDF cluster record (dmt.h, DMT_SDF_CLUSTER_FLOATS): nodes in post-order.
struct SdfNode {
    uint32_t op, shape, a, b;
    float q[4];
    float c[3];
};

void put_sdf(std::vector<float>& out, const std::vector<SdfNode>& nodes, const float lo[3], const float hi[3],
             uint32_t mat, const float alb[3], float rough, float ior) {
    float r[112] = {};
    auto u2f = [](uint32_t u) {
        float f;
        std::memcpy(&f, &u, 4);
        return f;
    };
    for (int a = 0; a < 3; ++a) {
        r[a] = lo[a];
        r[4 + a] = hi[a];
        r[8 + a] = alb[a];
    }
    r[3] = u2f((uint32_t)nodes.size());
    r[7] = u2f(mat);
    r[11] = rough;
    r[12] = ior;
    for (size_t i = 0; i < nodes.size(); ++i) {
        float* n = r + 16 + 12 * i;
        n[0] = u2f(nodes[i].op); n[1] = u2f(nodes[i].shape); n[2] = u2f(nodes[i].a); n[3] = u2f(nodes[i].b);
        for (int k = 0; k < 4; ++k) n[4 + k] = nodes[i].q[k];
        for (int k = 0; k < 3; ++k) n[8 + k] = nodes[i].c[k];
    }
    out.insert(out.end(), r, r + 112);
}

// SDF clusters of six kinds (i mod 6) with config 0 materials (seeded shuffle
// of i mod 3), centres uniform in [-extent, extent]^3; each box encloses the
// surface with a 1e-3 margin (smooth unions reach k / 4 past their parts).
void gen_sdf(Rng& rng, int n, std::vector<float>& out, float extent) {
    std::vector<uint32_t> mats((size_t)n);
    for (int i = 0; i < n; ++i) mats[(size_t)i] = (uint32_t)(i % 3);
    for (int i = n - 1; i > 0; --i) {
        int j = (int)(rng.next() % (uint64_t)(i + 1));
        std::swap(mats[(size_t)i], mats[(size_t)j]);
    }
    for (int i = 0; i < n; ++i) {
        const float c[3] = {rng.uf(-extent, extent), rng.uf(-extent, extent), rng.uf(-extent, extent)};
        float alb[3] = {rng.uf(0.1f, 0.9f), rng.uf(0.1f, 0.9f), rng.uf(0.1f, 0.9f)};
        float rough = rng.uf(0.0f, 0.5f);
        const uint32_t m = mats[(size_t)i];
        if (m == 0u) rough = 0.0f;
        if (m == 2u) { alb[0] = alb[1] = alb[2] = 1.0f; rough = 0.0f; }
        std::vector<SdfNode> nodes;
        float ext[3];   // half size of the box around c
        switch (i % 6) {
            case 0: {   // box
                const float h[3] = {rng.uf(0.08f, 0.25f), rng.uf(0.08f, 0.25f), rng.uf(0.08f, 0.25f)};
                nodes.push_back({0u, 1u, 0u, 0u, {h[0], h[1], h[2], 0.0f}, {c[0], c[1], c[2]}});
                for (int a = 0; a < 3; ++a) ext[a] = h[a];
                break;
            }
            case 1: {   // rounded box
                const float h[3] = {rng.uf(0.06f, 0.2f), rng.uf(0.06f, 0.2f), rng.uf(0.06f, 0.2f)}, rr = 0.04f;
                nodes.push_back({0u, 2u, 0u, 0u, {h[0], h[1], h[2], rr}, {c[0], c[1], c[2]}});
                for (int a = 0; a < 3; ++a) ext[a] = h[a] + rr;
                break;
            }
            case 2: {   // box united with a sphere (hard union)
                const float h[3] = {rng.uf(0.06f, 0.18f), rng.uf(0.06f, 0.18f), rng.uf(0.06f, 0.18f)};
                const float sr = rng.uf(0.08f, 0.2f);
                const float o[3] = {rng.uf(-0.15f, 0.15f), rng.uf(-0.15f, 0.15f), rng.uf(-0.15f, 0.15f)};
                nodes.push_back({0u, 1u, 0u, 0u, {h[0], h[1], h[2], 0.0f}, {c[0], c[1], c[2]}});
                nodes.push_back({0u, 0u, 0u, 0u, {sr, 0.0f, 0.0f, 0.0f}, {c[0] + o[0], c[1] + o[1], c[2] + o[2]}});
                nodes.push_back({1u, 0u, 0u, 1u, {0.0f, 0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}});
                for (int a = 0; a < 3; ++a) ext[a] = std::fmax(h[a], std::fabs(o[a]) + sr);
                break;
            }
            case 3: {   // sphere smoothly united with a capsule
                const float sr = rng.uf(0.08f, 0.18f), ch = rng.uf(0.05f, 0.15f), cr = rng.uf(0.04f, 0.1f), k = 0.08f;
                nodes.push_back({0u, 0u, 0u, 0u, {sr, 0.0f, 0.0f, 0.0f}, {c[0] - 0.12f, c[1], c[2]}});
                nodes.push_back({0u, 4u, 0u, 0u, {ch, cr, 0.0f, 0.0f}, {c[0] + 0.12f, c[1], c[2]}});
                nodes.push_back({1u, 0u, 0u, 1u, {k, 0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}});
                ext[0] = 0.12f + std::fmax(sr, cr) + 0.25f * k;
                ext[1] = std::fmax(sr, ch + cr) + 0.25f * k;
                ext[2] = std::fmax(sr, cr) + 0.25f * k;
                break;
            }
            case 4: {   // box with a sphere carved out of one corner (hard subtraction)
                const float h[3] = {rng.uf(0.1f, 0.22f), rng.uf(0.1f, 0.22f), rng.uf(0.1f, 0.22f)};
                const float sr = rng.uf(0.1f, 0.2f);
                nodes.push_back({0u, 1u, 0u, 0u, {h[0], h[1], h[2], 0.0f}, {c[0], c[1], c[2]}});
                nodes.push_back({0u, 0u, 0u, 0u, {sr, 0.0f, 0.0f, 0.0f}, {c[0] + h[0], c[1] + h[1], c[2] + h[2]}});
                nodes.push_back({2u, 0u, 0u, 1u, {0.0f, 0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}});
                for (int a = 0; a < 3; ++a) ext[a] = h[a];
                break;
            }
            default: {   // torus
                const float R = rng.uf(0.12f, 0.25f), r = rng.uf(0.03f, 0.07f);
                nodes.push_back({0u, 3u, 0u, 0u, {R, r, 0.0f, 0.0f}, {c[0], c[1], c[2]}});
                ext[0] = ext[2] = R + r;
                ext[1] = r;
                break;
            }
        }
        float lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = c[a] - (ext[a] + 1e-3f);
            hi[a] = c[a] + (ext[a] + 1e-3f);
        }
        put_sdf(out, nodes, lo, hi, m, alb, rough, m == 2u ? 1.5f : 1.0f);
    }
}

// 3-D value noise: trilinear blend (smoothstep weights) of hashed lattice values.
struct ValueNoise {
    uint64_t seed;
    float lattice(int x, int y, int z) const {
        uint64_t h = seed ^ (uint64_t)(uint32_t)x * 0x8DA6B343ull ^ (uint64_t)(uint32_t)y * 0xD8163841ull ^
                     (uint64_t)(uint32_t)z * 0xCB1AB31Full;
        h ^= h >> 33; h *= 0xFF51AFD7ED558CCDull; h ^= h >> 33; h *= 0xC4CEB9FE1A85EC53ull; h ^= h >> 33;
        return float(h >> 40) * (1.0f / 16777216.0f);
    }
    float operator()(float x, float y, float z) const {
        int ix = (int)std::floor(x), iy = (int)std::floor(y), iz = (int)std::floor(z);
        float fx = x - ix, fy = y - iy, fz = z - iz;
        fx = fx * fx * (3 - 2 * fx); fy = fy * fy * (3 - 2 * fy); fz = fz * fz * (3 - 2 * fz);
        float v = 0;
        for (int c = 0; c < 8; ++c) {
            int dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
            float w = (dx ? fx : 1 - fx) * (dy ? fy : 1 - fy) * (dz ? fz : 1 - fz);
            v += w * lattice(ix + dx, iy + dy, iz + dz);
        }
        return v;
    }
};

void icosphere(int subdiv, std::vector<float>& pos, std::vector<uint32_t>& idx) {
    const float t = (1.0f + std::sqrt(5.0f)) * 0.5f;
    const float base[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                               {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    const uint32_t faces[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                                   {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                                   {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    std::vector<double> P;
    for (auto& v : base) {
        double l = std::sqrt(double(v[0]) * v[0] + double(v[1]) * v[1] + double(v[2]) * v[2]);
        P.push_back(v[0] / l); P.push_back(v[1] / l); P.push_back(v[2] / l);
    }
    std::vector<uint32_t> F(&faces[0][0], &faces[0][0] + 60);
    for (int s = 0; s < subdiv; ++s) {
        std::unordered_map<uint64_t, uint32_t> mid;
        mid.reserve(F.size() * 2);
        auto midpoint = [&](uint32_t a, uint32_t b) {
            uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            double m[3] = {(P[3 * a] + P[3 * b]) * 0.5, (P[3 * a + 1] + P[3 * b + 1]) * 0.5,
                           (P[3 * a + 2] + P[3 * b + 2]) * 0.5};
            double l = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
            uint32_t id = (uint32_t)(P.size() / 3);
            P.push_back(m[0] / l); P.push_back(m[1] / l); P.push_back(m[2] / l);
            mid.emplace(key, id);
            return id;
        };
        std::vector<uint32_t> G;
        G.reserve(F.size() * 4);
        for (size_t f = 0; f < F.size(); f += 3) {
            uint32_t a = F[f], b = F[f + 1], c = F[f + 2];
            uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            uint32_t nf[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            G.insert(G.end(), nf, nf + 12);
        }
        F.swap(G);
    }
    pos.resize(P.size());
    for (size_t i = 0; i < P.size(); ++i) pos[i] = float(P[i]);
    idx = std::move(F);
}

}  // namespace

extern "C" {

dmt_status dmt_camera_look_at(const float eye[3], const float target[3], float fov_y_deg, uint32_t width,
                              uint32_t height, dmt_frame_params* f) {
    if (!eye || !target || !f || width == 0 || height == 0) return DMT_ERR_INVALID_ARG;
    float d[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
    float l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (!(l > 0.0f)) return DMT_ERR_INVALID_ARG;
    float fw[3] = {d[0] / l, d[1] / l, d[2] / l};
    // Camera::Right = normalize(cross(fwd, (0,1,0))); Camera::Up = normalize(cross(right, fwd)).
    float r[3] = {-fw[2], 0.0f, fw[0]};
    float rl = std::sqrt(r[0] * r[0] + r[2] * r[2]);
    if (!(rl > 0.0f)) return DMT_ERR_INVALID_ARG;
    r[0] /= rl; r[2] /= rl;
    float u[3] = {r[1] * fw[2] - r[2] * fw[1], r[2] * fw[0] - r[0] * fw[2], r[0] * fw[1] - r[1] * fw[0]};
    float ul = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    u[0] /= ul; u[1] /= ul; u[2] /= ul;
    f->pos_fovtan[0] = eye[0]; f->pos_fovtan[1] = eye[1]; f->pos_fovtan[2] = eye[2];
    f->pos_fovtan[3] = std::tan(fov_y_deg * 0.017453292519943295f * 0.5f);
    f->fwd_aspect[0] = fw[0]; f->fwd_aspect[1] = fw[1]; f->fwd_aspect[2] = fw[2];
    f->fwd_aspect[3] = float(width) / float(height);
    f->right_xyz[0] = r[0]; f->right_xyz[1] = r[1]; f->right_xyz[2] = r[2]; f->right_xyz[3] = 0.0f;
    f->up_xyz[0] = u[0]; f->up_xyz[1] = u[1]; f->up_xyz[2] = u[2]; f->up_xyz[3] = 0.0f;
    f->width = width;
    f->height = height;
    return DMT_OK;
}

dmt_status dmt_scene_generate(uint32_t config, uint64_t seed, const dmt_gen_options* opt, dmt_generated_scene** out) {
    if (!out) return DMT_ERR_INVALID_ARG;
    if (config > 4u || config == 3u) return DMT_ERR_UNSUPPORTED;
    dmt_gen_options o{-1, -1, -1, 0, 0, 0, 0, -1, -1};
    if (opt) o = *opt;
    Rng rng(seed + 0x1000u * config);
    std::vector<float> prims, lights, pos, env, sdf;
    std::vector<uint32_t> idx;
    uint32_t W = 256, H = 256, spp = 1, bounces = 4, ew = 8, eh = 4;
    const float zero3[3] = {0, 0, 0};
    if (config == 0u || config == 4u) {
        int n = o.spheres >= 0 ? o.spheres : (config == 0u ? 64 : 256);
        gen_spheres(rng, n, prims);
        if (config == 4u) { W = H = 512; spp = 16; }
    } else if (config == 2u) {
        // 3,584 spheres and 512 SDF clusters + a 500k-triangle mesh lit by
        // 4,096 emissive triangles
        W = 1280; H = 720; spp = 4; bounces = 6;
        int n = o.spheres >= 0 ? o.spheres : 3584;
        gen_spheres(rng, n, prims, 4.5f, 0.05f, 0.3f);
        int sub = o.subdivisions >= 0 ? o.subdivisions : 7;
        int nrand = o.random_tris >= 0 ? o.random_tris : 500000 - 20 * (1 << (2 * 7));
        icosphere(sub, pos, idx);
        for (auto& v : pos) v *= 2.5f;
        for (int t = 0; t < nrand; ++t) {
            float c[3] = {rng.uf(-5, 5), rng.uf(-5, 5), rng.uf(-5, 5)};
            uint32_t base = (uint32_t)(pos.size() / 3);
            for (int k = 0; k < 3; ++k) {
                float off[3] = {rng.uf(-1, 1), rng.uf(-1, 1), rng.uf(-1, 1)};
                float ol = std::sqrt(off[0] * off[0] + off[1] * off[1] + off[2] * off[2]);
                float s = ol > 0.0f ? 0.025f * rng.uf(0.2f, 1.0f) / ol : 0.0f;
                for (int a = 0; a < 3; ++a) pos.push_back(c[a] + off[a] * s);
            }
            idx.push_back(base); idx.push_back(base + 1); idx.push_back(base + 2);
        }
    } else {
        W = 1920; H = 1080; bounces = 8; ew = 512; eh = 256;
        int sub = o.subdivisions >= 0 ? o.subdivisions : 8;
        int nrand = o.random_tris >= 0 ? o.random_tris : 200000;
        icosphere(sub, pos, idx);
        ValueNoise vn{rng.next()};
        const float R = 3.0f, amp = 0.1f, freq = 4.0f;
        for (size_t v = 0; v < pos.size(); v += 3) {
            float nx = pos[v], ny = pos[v + 1], nz = pos[v + 2];
            float disp = amp * (2.0f * vn(nx * freq, ny * freq, nz * freq) - 1.0f);
            pos[v] = nx * (R + disp); pos[v + 1] = ny * (R + disp); pos[v + 2] = nz * (R + disp);
        }
        for (int t = 0; t < nrand; ++t) {
            float c[3] = {rng.uf(-5, 5), rng.uf(-5, 5), rng.uf(-5, 5)};
            uint32_t base = (uint32_t)(pos.size() / 3);
            for (int k = 0; k < 3; ++k) {
                // each vertex lies within 0.025 of the centre, so no edge
                // is longer than 0.05
                float off[3] = {rng.uf(-1, 1), rng.uf(-1, 1), rng.uf(-1, 1)};
                float ol = std::sqrt(off[0] * off[0] + off[1] * off[1] + off[2] * off[2]);
                float s = ol > 0.0f ? 0.025f * rng.uf(0.2f, 1.0f) / ol : 0.0f;
                for (int a = 0; a < 3; ++a) pos.push_back(c[a] + off[a] * s);
            }
            idx.push_back(base); idx.push_back(base + 1); idx.push_back(base + 2);
        }
    }
    if (o.width) W = o.width;
    if (o.height) H = o.height;
    if (o.env_width) ew = o.env_width;
    if (o.env_height) eh = o.env_height;
    env.assign((size_t)ew * eh * 3, 0.0f);
    if (config == 1u)
        for (auto& t : env) t = float(std::exp(rng.normal()));
    // config 1 with spheres: config 0's spheres and ground plane join the mesh
    // (drawn after everything else, so the default stream is unchanged)
    if (config == 1u && o.spheres > 0) gen_spheres(rng, o.spheres, prims);
    // config 4: every sphere also emits, emission U[0.02, 0.3] per channel, so
    // that all 7 material parameters of a sphere have a two-sided derivative
    // (a surface emits only when some channel is > 0)
    if (config == 4u)
        for (size_t i = 1; i < prims.size() / 28; ++i)
            for (int k = 0; k < 3; ++k) prims[i * 28 + 16 + k] = rng.uf(0.02f, 0.3f);
    if (config == 2u) {
        // 4,096 emissive triangles (edge ~0.3) facing down, in a slab above the
        // scene, emission exp(N(0,1)) per channel
        const int nl = o.lights >= 0 ? o.lights : 4096;
        for (int i = 0; i < nl; ++i) {
            float c[3] = {rng.uf(-6, 6), rng.uf(5, 7), rng.uf(-6, 6)};
            float v[3][3];
            for (int k = 0; k < 3; ++k)
                for (int a = 0; a < 3; ++a) v[k][a] = c[a] + rng.uf(-0.2f, 0.2f);
            const float e1[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
            const float e2[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
            const float ny = e1[2] * e2[0] - e1[0] * e2[2];
            if (ny > 0.0f)   // emit downwards: swap b and c
                for (int a = 0; a < 3; ++a) std::swap(v[1][a], v[2][a]);
            float L[20] = {};
            for (int a = 0; a < 3; ++a) {
                L[a] = v[0][a];
                L[4 + a] = float(std::exp(rng.normal()));
                L[8 + a] = v[1][a];
                L[12 + a] = v[2][a];
            }
            L[7] = 4.0f;
            L[19] = 1.0f;
            lights.insert(lights.end(), L, L + 20);
        }
    } else {
        const float lp[3] = {0.0f, 8.0f, 0.0f}, ln[3] = {0.0f, -1.0f, 0.0f}, lu[3] = {1.5f, 0.0f, 0.0f};
        const float I[3] = {10.0f, 10.0f, 10.0f};
        put_quad_light(lights, lp, ln, lu, 1.5f, I);
    }
    // config 2: the SDF clusters, drawn last
    if (config == 2u) gen_sdf(rng, o.sdf_clusters >= 0 ? o.sdf_clusters : 512, sdf, 4.5f);
    (void)zero3;
    auto* g = new (std::nothrow) dmt_generated_scene();
    if (!g) return DMT_ERR_OUT_OF_MEMORY;
    auto dup_f = [](const std::vector<float>& v) {
        float* p = v.empty() ? nullptr : (float*)std::malloc(v.size() * sizeof(float));
        if (p) std::memcpy(p, v.data(), v.size() * sizeof(float));
        return p;
    };
    g->prims = dup_f(prims); g->prim_count = (uint32_t)(prims.size() / 28);
    g->lights = dup_f(lights); g->light_count = (uint32_t)(lights.size() / 20);
    g->positions = dup_f(pos); g->vertex_count = (uint32_t)(pos.size() / 3);
    g->indices = idx.empty() ? nullptr : (uint32_t*)std::malloc(idx.size() * 4);
    if (g->indices) std::memcpy(g->indices, idx.data(), idx.size() * 4);
    g->index_count = (uint32_t)idx.size();
    g->env_rgb = dup_f(env); g->env_width = ew; g->env_height = eh;
    g->env_intensity = 1.0f;
    g->sdf_clusters = dup_f(sdf);
    g->sdf_cluster_count = (uint32_t)(sdf.size() / 112);
    std::memset(&g->frame, 0, sizeof(g->frame));
    const float eye[3] = {0.0f, 2.0f, 12.0f}, tgt[3] = {0.0f, 0.0f, 0.0f};
    dmt_camera_look_at(eye, tgt, 50.0f, W, H, &g->frame);
    g->frame.frame_index = 0;
    g->frame.spp = spp;
    g->frame.max_bounces = bounces;
    g->frame.reset_accum = 1;
    g->frame.mis_enabled = 1;
    g->frame.light_tree_enabled = 1;
    g->frame.refract_bounces = 4;
    g->frame.firefly_clamp = 10.0f;
    g->frame.ema_alpha = 0.0f;
    g->frame.sdf_max_iters = 128u;   // r_sdf_max_iters, r_sdf_epsilon defaults
    g->frame.sdf_epsilon = 1e-4f;
    *out = g;
    return DMT_OK;
}

dmt_status dmt_batch_workload_scene(uint64_t seed, uint32_t index, dmt_gen_options* opt, uint64_t* scene_seed,
                                    uint32_t* max_bounces) {
    if (!opt || !scene_seed || !max_bounces) return DMT_ERR_INVALID_ARG;
    Rng rng(seed * 0x2545F4914F6CDD1Dull + 0x3000u + index);
    static const uint32_t kRes[4] = {128u, 256u, 512u, 1024u};
    const uint32_t res = kRes[rng.next() % 4u];
    const double lt = std::log(1.0e4) + (std::log(2.0e6) - std::log(1.0e4)) * rng.u01();
    const uint32_t tris = (uint32_t)std::llround(std::exp(lt));
    int sub = 0;
    while (sub < 8 && 20u * (1u << (2 * (sub + 1))) <= tris) ++sub;
    const uint32_t ico = 20u * (1u << (2 * sub));
    std::memset(opt, 0, sizeof(*opt));
    opt->lights = -1;
    opt->sdf_clusters = -1;
    opt->subdivisions = sub;
    opt->random_tris = (int32_t)(tris - ico);
    opt->spheres = (index & 1u) ? 64 : 0;
    opt->width = res;
    opt->height = res;
    *scene_seed = rng.next();
    *max_bounces = 6u;
    return DMT_OK;
}

void dmt_scene_generated_free(dmt_generated_scene* g) {
    if (!g) return;
    std::free(g->prims); std::free(g->lights); std::free(g->positions); std::free(g->indices); std::free(g->env_rgb);
    std::free(g->sdf_clusters);
    delete g;
}

}  // extern "C"
