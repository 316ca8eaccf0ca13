// Host builders for the HIP path tracer.
//
// * build_bvh4: binned SAH (Wald 2007, "On fast construction of SAH-based
//   bounding volume hierarchies") over item bounds, then a top-down collapse
//   of the binary tree into 4-wide nodes (each node pulls up the children of
//   its largest-area internal child until it has four).  This replaces the
//   reference's centroid-median binary trees (src/renderer/TriangleBvh.cpp:33-169,
//   src/renderer/AnalyticBvh.cpp:369-486); the traversal result is the same
//   closest hit, the node layout is sized to a 128 B cache line.
// * build_light_tree: the reference light tree (src/renderer/LightTree.cpp:63-455),
//   same leaf summaries, cone union and median split, so the kernels draw
//   the same pick probabilities.
// * build_env_cdf: the reference env-map CDF (src/renderer/HdrImage.cpp:169-219).
// * extract_env_lights: the engine's HDRI light extraction at env-map load
//   (src/engine/Engine.cpp:4369-4532): bright-pixel threshold, connected
//   components, the brightest clusters as directional lights.
#include "prep.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <system_error>
#include <thread>

namespace dmt {
namespace {

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const float* l, const float* h) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], l[a]);
            hi[a] = std::max(hi[a], h[a]);
        }
    }
    void grow_pt(const float* p) { grow(p, p); }
    float area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0.0f)) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct BinNode {
    Box box;
    int32_t left = -1, right = -1;
    uint32_t first = 0, count = 0;   // leaf range in the index array
    bool leaf() const { return left < 0; }
};

constexpr uint32_t kMaxBins = 64;
// Subtrees at least this large are handed to another thread while spare
// threads remain; smaller ones are built where they are split.
constexpr uint32_t kParallelMin = 32768;

// One builder per thread: it owns the nodes of the subtrees it builds; the
// index array is shared, each builder touching only its own range.  The
// result does not depend on the thread count: a subtree's split choices see
// only its own items, and spliced node arrays keep every link.
struct SahBuilder {
    const BoundsItem* items;
    uint32_t* idx;
    std::vector<BinNode> nodes;
    uint32_t bins, max_leaf;
    std::atomic<int>* spare;

    SahBuilder(const BoundsItem* it, uint32_t* ix, uint32_t b, uint32_t ml, std::atomic<int>* sp)
        : items(it), idx(ix), bins(b), max_leaf(ml), spare(sp) {}

    // Chooses the split of idx[first, first + count); returns the partition
    // point, or first when the range becomes a leaf.
    uint32_t split(uint32_t first, uint32_t count, Box& box) {
        Box cbox;
        for (uint32_t i = first; i < first + count; ++i) {
            const BoundsItem& it = items[idx[i]];
            box.grow(it.lo, it.hi);
            cbox.grow_pt(it.c);
        }
        if (count <= 1) return first;
        const uint32_t B = bins;
        Box bb[3][kMaxBins];
        uint32_t bc[3][kMaxBins] = {};
        float scale[3];
        bool live[3];
        for (int a = 0; a < 3; ++a) {
            const float ext = cbox.hi[a] - cbox.lo[a];
            live[a] = ext > 0.0f;
            scale[a] = live[a] ? float(B) / ext : 0.0f;
        }
        // one pass over the items bins all three axes
        for (uint32_t i = first; i < first + count; ++i) {
            const BoundsItem& it = items[idx[i]];
            for (int a = 0; a < 3; ++a) {
                if (!live[a]) continue;
                const uint32_t k = std::min(B - 1, (uint32_t)((it.c[a] - cbox.lo[a]) * scale[a]));
                bb[a][k].grow(it.lo, it.hi);
                bc[a][k]++;
            }
        }
        float best_cost = INFINITY;
        int best_axis = -1;
        uint32_t best_split = 0;
        float right_area[kMaxBins];
        uint32_t right_count[kMaxBins];
        for (int a = 0; a < 3; ++a) {
            if (!live[a]) continue;
            Box acc;
            uint32_t n = 0;
            for (uint32_t k = B - 1; k > 0; --k) {
                acc.grow(bb[a][k].lo, bb[a][k].hi);
                n += bc[a][k];
                right_area[k] = acc.area();
                right_count[k] = n;
            }
            Box lacc;
            uint32_t ln = 0;
            for (uint32_t k = 0; k + 1 < B; ++k) {
                lacc.grow(bb[a][k].lo, bb[a][k].hi);
                ln += bc[a][k];
                if (ln == 0 || right_count[k + 1] == 0) continue;
                const float cost = lacc.area() * float(ln) + right_area[k + 1] * float(right_count[k + 1]);
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = a;
                    best_split = k + 1;
                }
            }
        }
        const float parent_area = box.area();
        const float leaf_cost = float(count);
        const float split_cost = parent_area > 0.0f ? 1.0f + best_cost / parent_area : INFINITY;
        if (count <= max_leaf && (best_axis < 0 || leaf_cost <= split_cost)) return first;
        if (best_axis < 0) return first + count / 2;   // coincident centroids: split by position
        const float lo = cbox.lo[best_axis], sc = scale[best_axis];
        const BoundsItem* its = items;
        const int ax = best_axis;
        uint32_t* it = std::partition(idx + first, idx + first + count, [&](uint32_t i) {
            return std::min(B - 1, (uint32_t)((its[i].c[ax] - lo) * sc)) < best_split;
        });
        uint32_t mid = (uint32_t)(it - idx);
        if (mid == first || mid == first + count) mid = first + count / 2;
        return mid;
    }

    int32_t build(uint32_t first, uint32_t count) {
        const int32_t me = (int32_t)nodes.size();
        nodes.emplace_back();
        Box box;
        const uint32_t mid = split(first, count, box);
        nodes[(size_t)me].box = box;
        if (mid == first) {
            nodes[(size_t)me].first = first;
            nodes[(size_t)me].count = count;
            return me;
        }
        int32_t l = -1, r = -1;
        const uint32_t nl = mid - first, nr = first + count - mid;
        bool forked = false;
        if (std::min(nl, nr) >= kParallelMin) {
            if (spare->fetch_sub(1) > 0) {
                SahBuilder sub(items, idx, bins, max_leaf, spare);
                int32_t sub_root = -1;
                std::thread th;
                try {
                    th = std::thread([&] { sub_root = sub.build(first, nl); });
                } catch (const std::system_error&) {
                    // no thread available: both halves are built here
                }
                if (th.joinable()) {
                    struct Join {
                        std::thread& t;
                        ~Join() { t.join(); }
                    } join{th};
                    r = build(mid, nr);
                }
                forked = sub_root >= 0;
                if (forked) {
                    const int32_t off = (int32_t)nodes.size();
                    for (BinNode n : sub.nodes) {
                        if (!n.leaf()) { n.left += off; n.right += off; }
                        nodes.push_back(n);
                    }
                    l = sub_root + off;
                }
            }
            spare->fetch_add(1);
        }
        if (!forked) {
            l = build(first, nl);
            r = build(mid, nr);
        }
        nodes[(size_t)me].left = l;
        nodes[(size_t)me].right = r;
        return me;
    }
};

struct Collapser {
    const std::vector<BinNode>& bn;
    std::vector<Node4>& out;
    uint32_t max_depth = 0;

    uint32_t child_ref(int32_t c, uint32_t depth) {
        const BinNode& n = bn[(size_t)c];
        if (n.leaf()) return leaf_ref(n.first, n.count);
        return collapse(c, depth + 1);
    }

    uint32_t collapse(int32_t root, uint32_t depth) {
        max_depth = std::max(max_depth, depth);
        int32_t kids[4];
        int nk = 0;
        const BinNode& r = bn[(size_t)root];
        if (r.leaf()) {
            kids[nk++] = root;
        } else {
            kids[nk++] = r.left;
            kids[nk++] = r.right;
            while (nk < 4) {
                int best = -1;
                float best_area = -1.0f;
                for (int i = 0; i < nk; ++i) {
                    const BinNode& k = bn[(size_t)kids[i]];
                    if (k.leaf()) continue;
                    float a = k.box.area();
                    if (a > best_area) { best_area = a; best = i; }
                }
                if (best < 0) break;
                int32_t e = kids[best];
                kids[best] = bn[(size_t)e].left;
                kids[nk++] = bn[(size_t)e].right;
            }
        }
        uint32_t me = (uint32_t)out.size();
        out.emplace_back();
        Node4 node{};
        for (int i = 0; i < 4; ++i) {
            node.lo_x[i] = node.lo_y[i] = node.lo_z[i] = 0.0f;
            node.hi_x[i] = node.hi_y[i] = node.hi_z[i] = 0.0f;
            node.child[i] = kEmptyChild;
            node.pad[i] = 0u;
        }
        for (int i = 0; i < nk; ++i) {
            const Box& b = bn[(size_t)kids[i]].box;
            node.lo_x[i] = b.lo[0]; node.hi_x[i] = b.hi[0];
            node.lo_y[i] = b.lo[1]; node.hi_y[i] = b.hi[1];
            node.lo_z[i] = b.lo[2]; node.hi_z[i] = b.hi[2];
        }
        for (int i = 0; i < nk; ++i) node.child[i] = child_ref(kids[i], depth);
        out[me] = node;
        return me;
    }
};

}  // namespace

Bvh4 build_bvh4(const std::vector<BoundsItem>& items, uint32_t bins, uint32_t max_leaf) {
    Bvh4 out;
    if (items.empty()) return out;
    if (bins < 2) bins = 16;
    bins = std::min(bins, kMaxBins);
    if (max_leaf < 1) max_leaf = 4;
    std::vector<uint32_t> idx(items.size());
    std::iota(idx.begin(), idx.end(), 0u);
    const unsigned hw = std::thread::hardware_concurrency();
    std::atomic<int> spare{(int)std::min(hw ? hw : 1u, 64u) - 1};
    SahBuilder b(items.data(), idx.data(), bins, max_leaf, &spare);
    b.nodes.reserve(items.size() * 2);
    b.build(0, (uint32_t)items.size());
    Collapser c{b.nodes, out.nodes};
    out.nodes.reserve(b.nodes.size() / 2 + 1);
    c.collapse(0, 1);
    out.depth = c.max_depth;
    out.order = std::move(idx);
    return out;
}

// ---------------------------------------------------------------------------
// Light tree (LightTree.cpp).  Node floats: [0..2] aabb_min, [3] left_first
// bits, [4..6] aabb_max, [7] count bits, [8..10] cone axis, [11] cos_half,
// [12] intensity, [13..15] 0.

namespace {

struct LtLeaf {
    float lo[3], hi[3], axis[3];
    float cos_half, power;
    uint32_t light;
};

// The engine rotates the spot axis / quad normal and the quad u vector by the
// light's orientation before summarising (Engine.cpp:5479-5508).
void orient_rotate(const float* q, const float* v, float* out) {
    const float cx = q[1] * v[2] - q[2] * v[1] + q[3] * v[0];
    const float cy = q[2] * v[0] - q[0] * v[2] + q[3] * v[1];
    const float cz = q[0] * v[1] - q[1] * v[0] + q[3] * v[2];
    out[0] = v[0] + 2.0f * (q[1] * cz - q[2] * cy);
    out[1] = v[1] + 2.0f * (q[2] * cx - q[0] * cz);
    out[2] = v[2] + 2.0f * (q[0] * cy - q[1] * cx);
}

// Emissive triangle (light type 4, no counterpart in the reference's light
// types): vertices a = v0.xyz, b = v2.xyz, c = v3.xyz in world space (no
// orientation), one-sided along normalize(cross(b - a, c - a)).
LtLeaf lt_leaf_triangle(const float* L, uint32_t id) {
    LtLeaf s{};
    s.light = id;
    const float* v[3] = {&L[0], &L[8], &L[12]};
    for (int a = 0; a < 3; ++a) {
        s.lo[a] = std::min(v[0][a], std::min(v[1][a], v[2][a]));
        s.hi[a] = std::max(v[0][a], std::max(v[1][a], v[2][a]));
    }
    const float e1[3] = {L[8] - L[0], L[9] - L[1], L[10] - L[2]};
    const float e2[3] = {L[12] - L[0], L[13] - L[1], L[14] - L[2]};
    const float c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const float len = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    for (int a = 0; a < 3; ++a) s.axis[a] = c[a] / len;
    s.cos_half = 0.0f;
    s.power = (0.2126f * L[4] + 0.7152f * L[5] + 0.0722f * L[6]) * (0.5f * len);
    return s;
}

LtLeaf lt_leaf(const float* Lp, uint32_t id) {
    if ((uint32_t)(Lp[7] + 0.5f) == 4u) return lt_leaf_triangle(Lp, id);
    float L[20];
    std::memcpy(L, Lp, sizeof(L));
    orient_rotate(&Lp[16], &Lp[8], &L[8]);
    if ((uint32_t)(Lp[7] + 0.5f) != 1u) orient_rotate(&Lp[16], &Lp[12], &L[12]);
    LtLeaf s{};
    s.light = id;
    const uint32_t type = (uint32_t)(L[7] + 0.5f);
    const bool point_like = type == 0u || type == 1u;
    if (point_like) {
        for (int a = 0; a < 3; ++a) { s.lo[a] = L[a] - 1e-3f; s.hi[a] = L[a] + 1e-3f; }
    } else if (type == 2u) {
        const float r = std::max(L[3], 1e-3f);
        for (int a = 0; a < 3; ++a) { s.lo[a] = L[a] - r; s.hi[a] = L[a] + r; }
    } else {
        const float vh = std::fabs(L[15]);
        for (int a = 0; a < 3; ++a) {
            const float e = std::fabs(L[12 + a]) + vh;
            s.lo[a] = L[a] - e;
            s.hi[a] = L[a] + e;
        }
    }
    if (type == 0u || type == 2u) {
        s.axis[0] = 0.0f; s.axis[1] = 1.0f; s.axis[2] = 0.0f;
        s.cos_half = -1.0f;
    } else {
        s.axis[0] = L[8]; s.axis[1] = L[9]; s.axis[2] = L[10];
        s.cos_half = (type == 1u) ? L[11] : 0.0f;
    }
    float p = 0.2126f * L[4] + 0.7152f * L[5] + 0.0722f * L[6];
    if (type == 2u) {
        const float r = std::max(L[3], 1e-3f);
        p *= 4.0f * 3.14159265358979f * r * r;
    } else if (type == 3u) {
        const float ul = std::sqrt(L[12] * L[12] + L[13] * L[13] + L[14] * L[14]);
        p *= 4.0f * ul * std::fabs(L[15]);
    }
    s.power = p;
    return s;
}

// Bounding cone of two cones (LightTree.cpp:171-273).
void cone_union(const float* a, float ca, const float* b, float cb, float* axis, float* cos_half) {
    auto copy_a = [&](float c) {
        axis[0] = a[0]; axis[1] = a[1]; axis[2] = a[2];
        *cos_half = c;
    };
    if (ca <= -0.999f || cb <= -0.999f) { copy_a(-1.0f); return; }
    const float d = std::clamp(a[0] * b[0] + a[1] * b[1] + a[2] * b[2], -1.0f, 1.0f);
    const float th_d = std::acos(d);
    const float th_a = std::acos(std::clamp(ca, -1.0f, 1.0f));
    const float th_b = std::acos(std::clamp(cb, -1.0f, 1.0f));
    if (th_d + th_b <= th_a) { copy_a(ca); return; }
    if (th_d + th_a <= th_b) {
        axis[0] = b[0]; axis[1] = b[1]; axis[2] = b[2];
        *cos_half = cb;
        return;
    }
    const float th_c = 0.5f * (th_a + th_b + th_d);
    if (th_c >= 3.14159265358979f) { copy_a(-1.0f); return; }
    *cos_half = std::cos(th_c);
    const float rot = th_c - th_a;
    const float cr = std::cos(rot), sr = std::sin(rot);
    float perp[3] = {b[0] - d * a[0], b[1] - d * a[1], b[2] - d * a[2]};
    const float pl = std::sqrt(perp[0] * perp[0] + perp[1] * perp[1] + perp[2] * perp[2]);
    if (!(pl > 1e-6f)) {
        axis[0] = a[0]; axis[1] = a[1]; axis[2] = a[2];
        if (d < 0.0f) *cos_half = -1.0f;
        return;
    }
    const float ip = 1.0f / pl;
    for (float& v : perp) v *= ip;
    for (int k = 0; k < 3; ++k) axis[k] = cr * a[k] + sr * perp[k];
    const float al = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    if (al > 1e-6f) {
        const float ia = 1.0f / al;
        for (int k = 0; k < 3; ++k) axis[k] *= ia;
    }
}

void lt_build(std::vector<LtLeaf>& lv, std::vector<float>& nodes, uint32_t b, uint32_t e, uint32_t dst) {
    float* me = &nodes[(size_t)dst * 16];
    if (e - b == 1) {
        const LtLeaf& s = lv[b];
        for (int a = 0; a < 3; ++a) { me[a] = s.lo[a]; me[4 + a] = s.hi[a]; me[8 + a] = s.axis[a]; }
        uint32_t one = 1u;
        std::memcpy(&me[3], &s.light, 4);
        std::memcpy(&me[7], &one, 4);
        me[11] = s.cos_half;
        me[12] = s.power;
        return;
    }
    float cl[3] = {INFINITY, INFINITY, INFINITY}, ch[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = b; i < e; ++i)
        for (int a = 0; a < 3; ++a) {
            const float c = 0.5f * (lv[i].lo[a] + lv[i].hi[a]);
            cl[a] = std::min(cl[a], c);
            ch[a] = std::max(ch[a], c);
        }
    int axis = 0;
    for (int a = 1; a < 3; ++a)
        if (ch[a] - cl[a] > ch[axis] - cl[axis]) axis = a;
    const uint32_t mid = b + (e - b) / 2;
    if (!(ch[axis] - cl[axis] < 1e-6f)) {
        std::nth_element(lv.begin() + b, lv.begin() + mid, lv.begin() + e, [axis](const LtLeaf& x, const LtLeaf& y) {
            return 0.5f * (x.lo[axis] + x.hi[axis]) < 0.5f * (y.lo[axis] + y.hi[axis]);
        });
    }
    const uint32_t left = (uint32_t)(nodes.size() / 16);
    nodes.resize(nodes.size() + 32, 0.0f);
    lt_build(lv, nodes, b, mid, left);
    lt_build(lv, nodes, mid, e, left + 1);
    me = &nodes[(size_t)dst * 16];
    const float* L = &nodes[(size_t)left * 16];
    const float* R = &nodes[(size_t)(left + 1) * 16];
    for (int a = 0; a < 3; ++a) {
        me[a] = std::min(L[a], R[a]);
        me[4 + a] = std::max(L[4 + a], R[4 + a]);
    }
    me[12] = L[12] + R[12];
    cone_union(&L[8], L[11], &R[8], R[11], &me[8], &me[11]);
    uint32_t zero = 0u;
    std::memcpy(&me[3], &left, 4);
    std::memcpy(&me[7], &zero, 4);
}

}  // namespace

std::vector<float> build_light_tree(const float* lights, uint32_t n) {
    std::vector<float> nodes;
    if (n == 0) return nodes;
    std::vector<LtLeaf> lv;
    lv.reserve(n);
    for (uint32_t i = 0; i < n; ++i) lv.push_back(lt_leaf(lights + (size_t)20 * i, i));
    nodes.reserve((size_t)32 * n);
    nodes.resize(16, 0.0f);
    lt_build(lv, nodes, 0, n, 0);
    return nodes;
}

double build_env_cdf(const float* rgb, uint32_t W, uint32_t H, std::vector<float>& marginal,
                     std::vector<float>& conditional, const uint8_t* mask) {
    marginal.assign(H, 0.0f);
    conditional.assign((size_t)W * H, 0.0f);
    double total = 0.0;
    for (uint32_t v = 0; v < H; ++v) {
        const double s = std::sin(3.14159265358979323846 * (double(v) + 0.5) / double(H));
        double run = 0.0;
        float* row = &conditional[(size_t)v * W];
        const float* px = rgb + (size_t)v * W * 3;
        for (uint32_t u = 0; u < W; ++u) {
            float lum = 0.2126f * px[3 * u] + 0.7152f * px[3 * u + 1] + 0.0722f * px[3 * u + 2];
            if (mask && mask[(size_t)v * W + u]) lum = 0.0f;   // an extracted cluster's pixel
            run += double(lum) * s;
            row[u] = float(run);
        }
        const double inv = run > 0.0 ? 1.0 / run : 0.0;
        for (uint32_t u = 0; u < W; ++u) row[u] = float(double(row[u]) * inv);
        marginal[v] = float(run);
        total += run;
    }
    const double inv = total > 0.0 ? 1.0 / total : 0.0;
    double run = 0.0;
    for (uint32_t v = 0; v < H; ++v) {
        run += marginal[v];
        marginal[v] = float(run * inv);
    }
    if (H > 0 && total > 0.0) marginal[H - 1] = 1.0f;
    return total;
}

uint32_t extract_env_lights(const float* rgb, uint32_t W, uint32_t H, double top_frac, float dir[][4],
                            float col[][4], uint8_t* mask) {
    const size_t N = (size_t)W * H;
    std::fill(mask, mask + N, uint8_t(0));
    if (N == 0) return 0u;
    // bright pixels: luminance at or above the top-fraction quantile, floored
    // at 5 % of the peak and at 0.5; nothing on an image whose peak is <= 0.5
    std::vector<float> lum(N);
    float peak = 0.0f;
    for (size_t p = 0; p < N; ++p) {
        lum[p] = 0.2126f * rgb[3 * p] + 0.7152f * rgb[3 * p + 1] + 0.0722f * rgb[3 * p + 2];
        peak = std::max(peak, lum[p]);
    }
    if (!(peak > 0.5f)) return 0u;
    top_frac = std::min(std::max(top_frac, 1e-6), 0.5);
    const size_t k = N - std::max<size_t>(1, size_t(double(N) * top_frac));
    std::vector<float> order_stat = lum;
    std::nth_element(order_stat.begin(), order_stat.begin() + (std::ptrdiff_t)k, order_stat.end());
    float thresh = order_stat[k];
    if (thresh < 0.05f * peak) thresh = 0.05f * peak;
    if (thresh < 0.5f) thresh = 0.5f;
    for (size_t p = 0; p < N; ++p) mask[p] = lum[p] >= thresh ? 1u : 0u;

    // 4-connected components of the bright pixels (columns wrap around, rows
    // do not), each walked depth-first from its first pixel in row-major
    // order; per component the luminance-weighted direction sum, the flux and
    // the luminance integral over the solid angle
    struct Cluster {
        double dir[3] = {0.0, 0.0, 0.0};
        double wsum = 0.0;
        double flux[3] = {0.0, 0.0, 0.0};
        double lum = 0.0;
        uint32_t pixels = 0;
    };
    const double kPi = 3.14159265358979323846;
    const double d_omega = (2.0 * kPi * kPi) / double(N);
    std::vector<Cluster> cl;
    std::vector<uint32_t> id(N, 0u);   // 1-based component, 0 = none yet
    std::vector<std::pair<uint32_t, uint32_t>> todo;
    for (uint32_t v0 = 0; v0 < H; ++v0) {
        for (uint32_t u0 = 0; u0 < W; ++u0) {
            const size_t p0 = (size_t)v0 * W + u0;
            if (!mask[p0] || id[p0] != 0u) continue;
            cl.emplace_back();
            const uint32_t cid = (uint32_t)cl.size();
            Cluster& c = cl.back();
            id[p0] = cid;
            todo.assign(1, {u0, v0});
            while (!todo.empty()) {
                const auto [u, v] = todo.back();
                todo.pop_back();
                const size_t p = (size_t)v * W + u;
                const double theta = kPi * (double(v) + 0.5) / double(H);
                const double st = std::sin(theta), ct = std::cos(theta);
                const double dw = d_omega * st;
                const double phi = 2.0 * kPi * (double(u) + 0.5) / double(W) - kPi;
                const float r = rgb[3 * p], g = rgb[3 * p + 1], b = rgb[3 * p + 2];
                const double l = 0.2126 * r + 0.7152 * g + 0.0722 * b;
                const double wl = l * dw;
                c.dir[0] += (st * std::cos(phi)) * wl;
                c.dir[1] += ct * wl;
                c.dir[2] += (st * std::sin(phi)) * wl;
                c.wsum += wl;
                c.flux[0] += double(r) * dw;
                c.flux[1] += double(g) * dw;
                c.flux[2] += double(b) * dw;
                c.lum += l * dw;
                c.pixels += 1u;
                const uint32_t nb[4][2] = {{u == 0u ? W - 1u : u - 1u, v}, {u + 1u >= W ? 0u : u + 1u, v},
                                           {u, v - 1u}, {u, v + 1u}};
                for (int k2 = 0; k2 < 4; ++k2) {
                    if ((k2 == 2 && v == 0u) || (k2 == 3 && v + 1u >= H)) continue;
                    const size_t q = (size_t)nb[k2][1] * W + nb[k2][0];
                    if (mask[q] && id[q] == 0u) {
                        id[q] = cid;
                        todo.push_back({nb[k2][0], nb[k2][1]});
                    }
                }
            }
        }
    }
    // the brightest (luminance integral) kMaxHdriLights components of at
    // least 3 pixels become lights, picked in proportion to that integral;
    // the other components go back into the importance CDF
    std::vector<uint32_t> rank(cl.size());
    std::iota(rank.begin(), rank.end(), 0u);
    std::sort(rank.begin(), rank.end(), [&](uint32_t a, uint32_t b) { return cl[a].lum > cl[b].lum; });
    std::vector<uint8_t> keep(cl.size() + 1u, 0u);
    uint32_t n = 0;
    float lum_f[kMaxEnvLights] = {};
    double total = 0.0;
    for (uint32_t i : rank) {
        if (n >= kMaxEnvLights) break;
        const Cluster& c = cl[i];
        if (c.pixels < 3u || !(c.wsum > 0.0)) continue;
        keep[i + 1u] = 1u;
        const double inv = 1.0 / std::sqrt((c.dir[0] * c.dir[0] + c.dir[1] * c.dir[1]) + c.dir[2] * c.dir[2]);
        for (int a = 0; a < 3; ++a) {
            dir[n][a] = float(c.dir[a] * inv);
            col[n][a] = float(c.flux[a]);
        }
        col[n][3] = 0.0f;
        lum_f[n] = float(c.lum);
        total += c.lum;
        n++;
    }
    if (total > 0.0) {
        const float inv_total = float(1.0 / total);
        for (uint32_t i = 0; i < n; ++i) dir[i][3] = lum_f[i] * inv_total;
    }
    for (size_t p = 0; p < N; ++p)
        if (id[p] != 0u && !keep[id[p]]) mask[p] = 0u;
    return n;
}

}  // namespace dmt
