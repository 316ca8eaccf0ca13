// Device BVH4 build: see lbvh.h.
#include "lbvh.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "dmt_dev.h"

namespace dmt {

namespace {

constexpr uint32_t kBinLeaf = 0x80000000u;   // radix-tree child reference: leaf (sorted position)

uint32_t blocks(uint32_t n) { return std::max(1u, std::min((n + 255u) / 256u, 1u << 16)); }

__device__ __forceinline__ f3 vtx(const float* p, uint32_t i) { return f3{p[3u * i], p[3u * i + 1u], p[3u * i + 2u]}; }

// 10 bits of v spread to every third bit.
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3FFu;
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__device__ __forceinline__ uint32_t quantize(float c, float lo, float scale) {
    const float q = (c - lo) * scale;
    return q <= 0.0f ? 0u : (q >= 1023.0f ? 1023u : (uint32_t)q);
}

__global__ __launch_bounds__(256) void k_morton(const float* __restrict__ pos, const uint32_t* __restrict__ idx,
                                                uint32_t n, f3 lo, f3 scale, uint32_t* __restrict__ keys,
                                                uint32_t* __restrict__ vals) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
        const f3 a = vtx(pos, idx[3u * t]), b = vtx(pos, idx[3u * t + 1u]), c = vtx(pos, idx[3u * t + 2u]);
        const f3 m = ((a + b) + c) * (1.0f / 3.0f);
        keys[t] = (spread10(quantize(m.x, lo.x, scale.x)) << 2) | (spread10(quantize(m.y, lo.y, scale.y)) << 1) |
                  spread10(quantize(m.z, lo.z, scale.z));
        vals[t] = t;
    }
}

// Vertex bounds of the triangle at each sorted position, and its record.
__global__ __launch_bounds__(256) void k_leaves(const float* __restrict__ pos, const uint32_t* __restrict__ idx,
                                                const uint32_t* __restrict__ order, uint32_t n,
                                                float4* __restrict__ leaf_lo, float4* __restrict__ leaf_hi,
                                                float4* __restrict__ tris, uint32_t* __restrict__ leaf_idx) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const uint32_t t = order[k];
        const uint32_t ia = idx[3u * t], ib = idx[3u * t + 1u], ic = idx[3u * t + 2u];
        const f3 a = vtx(pos, ia), b = vtx(pos, ib), c = vtx(pos, ic);
        // the host build's item bounds: fmin(v0, fmin(v1, v2)), fmax likewise
        leaf_lo[k] = make_float4(fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)), fminf(a.z, fminf(b.z, c.z)), 0.0f);
        leaf_hi[k] = make_float4(fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)), fmaxf(a.z, fmaxf(b.z, c.z)), 0.0f);
        float4* r = tris + 3u * (size_t)k;
        r[0] = make_float4(a.x, a.y, a.z, __uint_as_float(t));
        r[1] = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.0f);
        r[2] = make_float4(c.x - a.x, c.y - a.y, c.z - a.z, 0.0f);
        leaf_idx[3u * k] = ia;
        leaf_idx[3u * k + 1u] = ib;
        leaf_idx[3u * k + 2u] = ic;
    }
}

// Common-prefix length of sorted keys i and j (-1 outside 0..n-1); equal keys
// are told apart by their positions.
__device__ __forceinline__ int prefix(const uint32_t* keys, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    const uint32_t a = keys[i], b = keys[j];
    if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clz(a ^ b);
}

// Internal node i of the radix tree: its key range, split and children.
__global__ __launch_bounds__(256) void k_tree(const uint32_t* __restrict__ keys, uint32_t n, uint32_t* __restrict__ left,
                                              uint32_t* __restrict__ right, uint2* __restrict__ range,
                                              int32_t* __restrict__ parent_int, int32_t* __restrict__ parent_leaf) {
    const uint32_t m = n - 1u;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t ii = blockIdx.x * blockDim.x + threadIdx.x; ii < m; ii += stride) {
        const int64_t i = ii, N = n;
        const int64_t d = (prefix(keys, N, i, i + 1) - prefix(keys, N, i, i - 1)) > 0 ? 1 : -1;
        const int dmin = prefix(keys, N, i, i - d);
        int64_t lmax = 2;
        while (prefix(keys, N, i, i + lmax * d) > dmin) lmax *= 2;
        int64_t l = 0;
        for (int64_t t = lmax / 2; t >= 1; t /= 2)
            if (prefix(keys, N, i, i + (l + t) * d) > dmin) l += t;
        const int64_t j = i + l * d;
        const int dnode = prefix(keys, N, i, j);
        int64_t s = 0, t = l;
        do {
            t = (t + 1) / 2;
            if (prefix(keys, N, i, i + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        const int64_t g = i + s * d + (d < 0 ? -1 : 0);
        const int64_t a = i < j ? i : j, b = i < j ? j : i;
        range[ii] = make_uint2((uint32_t)a, (uint32_t)b);
        if (a == g) {
            left[ii] = kBinLeaf | (uint32_t)g;
            parent_leaf[g] = (int32_t)ii;
        } else {
            left[ii] = (uint32_t)g;
            parent_int[g] = (int32_t)ii;
        }
        if (b == g + 1) {
            right[ii] = kBinLeaf | (uint32_t)(g + 1);
            parent_leaf[g + 1] = (int32_t)ii;
        } else {
            right[ii] = (uint32_t)(g + 1);
            parent_int[g + 1] = (int32_t)ii;
        }
    }
}

__device__ __forceinline__ float aload(const float* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void astore(float* p, float v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bottom-up boxes: every leaf walks towards the root; at each internal node
// the second thread to arrive forms the union of both children.  Boxes of
// internal nodes are written and read with device-coherent accesses.
__global__ __launch_bounds__(256) void k_boxes(uint32_t n, const uint32_t* __restrict__ left,
                                               const uint32_t* __restrict__ right,
                                               const int32_t* __restrict__ parent_int,
                                               const int32_t* __restrict__ parent_leaf,
                                               const float4* __restrict__ leaf_lo, const float4* __restrict__ leaf_hi,
                                               float* node_box, uint32_t* arrivals) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        int32_t p = parent_leaf[k];
        while (p >= 0) {
            __threadfence();
            if (atomicAdd(arrivals + p, 1u) == 0u) break;   // the sibling's walk finishes this node
            __threadfence();
            float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
            const uint32_t ch[2] = {left[p], right[p]};
            for (int c = 0; c < 2; ++c) {
                float v[6];
                if (ch[c] & kBinLeaf) {
                    const float4 lo = leaf_lo[ch[c] & ~kBinLeaf], hi = leaf_hi[ch[c] & ~kBinLeaf];
                    v[0] = lo.x; v[1] = lo.y; v[2] = lo.z; v[3] = hi.x; v[4] = hi.y; v[5] = hi.z;
                } else {
                    for (int a = 0; a < 6; ++a) v[a] = aload(node_box + 6u * (size_t)ch[c] + a);
                }
                for (int a = 0; a < 3; ++a) {
                    b[a] = fminf(b[a], v[a]);
                    b[3 + a] = fmaxf(b[3 + a], v[3 + a]);
                }
            }
            for (int a = 0; a < 6; ++a) astore(node_box + 6u * (size_t)p + a, b[a]);
            p = parent_int[p];
        }
    }
}

__device__ __forceinline__ void box_of(uint32_t ref, const float* node_box, const float4* leaf_lo,
                                       const float4* leaf_hi, float v[6]) {
    if (ref & kBinLeaf) {
        const float4 lo = leaf_lo[ref & ~kBinLeaf], hi = leaf_hi[ref & ~kBinLeaf];
        v[0] = lo.x; v[1] = lo.y; v[2] = lo.z; v[3] = hi.x; v[4] = hi.y; v[5] = hi.z;
    } else {
        for (int a = 0; a < 6; ++a) v[a] = node_box[6u * (size_t)ref + a];
    }
}

// One level of the collapse: each queued radix-tree node becomes the 4-wide
// node `entry.y`; its internal children with more than max_leaf triangles
// get the next node indices and go to the next level's queue.
__global__ __launch_bounds__(256) void k_collapse(const uint2* __restrict__ q_in, uint32_t count,
                                                  const uint32_t* __restrict__ left, const uint32_t* __restrict__ right,
                                                  const uint2* __restrict__ range, const float* __restrict__ node_box,
                                                  const float4* __restrict__ leaf_lo, const float4* __restrict__ leaf_hi,
                                                  uint32_t max_leaf, float4* __restrict__ nodes, uint2* __restrict__ q_out,
                                                  uint32_t* __restrict__ counters) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < count; e += stride) {
        const uint2 entry = q_in[e];
        uint32_t kids[4] = {left[entry.x], right[entry.x], 0u, 0u};
        uint32_t nk = 2u;
        auto size = [&](uint32_t r) { return (r & kBinLeaf) ? 1u : range[r].y - range[r].x + 1u; };
        while (nk < 4u) {
            int best = -1;
            float best_area = -1.0f;
            for (uint32_t q = 0; q < nk; ++q) {
                if ((kids[q] & kBinLeaf) || size(kids[q]) <= max_leaf) continue;
                float v[6];
                box_of(kids[q], node_box, leaf_lo, leaf_hi, v);
                const float dx = v[3] - v[0], dy = v[4] - v[1], dz = v[5] - v[2];
                const float area = 2.0f * (dx * dy + dy * dz + dz * dx);
                if (area > best_area) {
                    best_area = area;
                    best = (int)q;
                }
            }
            if (best < 0) break;
            const uint32_t open = kids[best];
            kids[best] = left[open];
            kids[nk++] = right[open];
        }
        float row[6][4];
        uint32_t ref[4];
        for (uint32_t q = 0; q < 4u; ++q) {
            if (q >= nk) {
                for (int a = 0; a < 6; ++a) row[a][q] = 0.0f;
                ref[q] = kEmpty;
                continue;
            }
            float v[6];
            box_of(kids[q], node_box, leaf_lo, leaf_hi, v);
            row[0][q] = v[0]; row[1][q] = v[3];
            row[2][q] = v[1]; row[3][q] = v[4];
            row[4][q] = v[2]; row[5][q] = v[5];
            const uint32_t sz = size(kids[q]);
            if (sz <= max_leaf) {
                const uint32_t first = (kids[q] & kBinLeaf) ? (kids[q] & ~kBinLeaf) : range[kids[q]].x;
                ref[q] = kLeaf | ((sz - 1u) << 27) | first;
            } else {
                const uint32_t idx = atomicAdd(counters, 1u);
                q_out[atomicAdd(counters + 1, 1u)] = make_uint2(kids[q], idx);
                ref[q] = idx;
            }
        }
        float4* out = nodes + 8u * (size_t)entry.y;
        for (int a = 0; a < 6; ++a) out[a] = make_float4(row[a][0], row[a][1], row[a][2], row[a][3]);
        out[6] = make_float4(__uint_as_float(ref[0]), __uint_as_float(ref[1]), __uint_as_float(ref[2]),
                             __uint_as_float(ref[3]));
        out[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// A one-triangle mesh: the root holds one leaf.
__global__ void k_single_root(const float4* leaf_lo, const float4* leaf_hi, float4* nodes) {
    const float4 lo = leaf_lo[0], hi = leaf_hi[0];
    nodes[0] = make_float4(lo.x, 0.0f, 0.0f, 0.0f);
    nodes[1] = make_float4(hi.x, 0.0f, 0.0f, 0.0f);
    nodes[2] = make_float4(lo.y, 0.0f, 0.0f, 0.0f);
    nodes[3] = make_float4(hi.y, 0.0f, 0.0f, 0.0f);
    nodes[4] = make_float4(lo.z, 0.0f, 0.0f, 0.0f);
    nodes[5] = make_float4(hi.z, 0.0f, 0.0f, 0.0f);
    nodes[6] = make_float4(__uint_as_float(kLeaf), __uint_as_float(kEmpty), __uint_as_float(kEmpty),
                           __uint_as_float(kEmpty));
    nodes[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Temporary device allocations, freed when the build returns.
struct Scratch {
    std::vector<void*> ptrs;
    template <class T> hipError_t get(T** p, size_t count) {
        *p = nullptr;
        hipError_t e = hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) ptrs.push_back(*p);
        else *p = nullptr;
        return e;
    }
    ~Scratch() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

}  // namespace

hipError_t build_lbvh(const LbvhInput& in, uint32_t tri_pad, LbvhOutput& out, hipStream_t stream) {
    const uint32_t n = in.tri_count;
    if (n == 0) return hipErrorInvalidValue;
    const uint32_t m = n > 1u ? n - 1u : 0u;
    Scratch tmp;
    uint32_t *keys = nullptr, *vals = nullptr, *keys2 = nullptr, *order = nullptr;
    float4 *leaf_lo = nullptr, *leaf_hi = nullptr;
    hipError_t e = tmp.get(&keys, n);
    if (e == hipSuccess) e = tmp.get(&vals, n);
    if (e == hipSuccess) e = tmp.get(&keys2, n);
    if (e == hipSuccess) e = tmp.get(&order, n);
    if (e == hipSuccess) e = tmp.get(&leaf_lo, n);
    if (e == hipSuccess) e = tmp.get(&leaf_hi, n);
    if (e != hipSuccess) return e;
    // outputs
    out = LbvhOutput{};
    auto fail = [&out](hipError_t err) {
        (void)hipFree(out.tris);
        (void)hipFree(out.leaf_idx);
        (void)hipFree(out.nodes);
        out.tris = nullptr;
        out.leaf_idx = nullptr;
        out.nodes = nullptr;
        return err;
    };
    if ((e = hipMalloc((void**)&out.tris, ((size_t)n * 3u + tri_pad) * sizeof(float4))) != hipSuccess) return fail(e);
    if ((e = hipMemsetAsync(out.tris + (size_t)n * 3u, 0, tri_pad * sizeof(float4), stream)) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void**)&out.leaf_idx, (size_t)n * 3u * sizeof(uint32_t))) != hipSuccess) return fail(e);

    // Morton order
    const f3 lo{in.cmin[0], in.cmin[1], in.cmin[2]};
    f3 scale;
    scale.x = 1024.0f / std::max(in.cmax[0] - in.cmin[0], 1e-30f);
    scale.y = 1024.0f / std::max(in.cmax[1] - in.cmin[1], 1e-30f);
    scale.z = 1024.0f / std::max(in.cmax[2] - in.cmin[2], 1e-30f);
    hipLaunchKernelGGL(k_morton, dim3(blocks(n)), dim3(256), 0, stream, in.positions, in.indices, n, lo, scale, keys, vals);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    size_t sort_bytes = 0;
    if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys, keys2, vals, order, (int)n, 0, 30,
                                                stream)) != hipSuccess)
        return fail(e);
    uint8_t* sort_tmp = nullptr;
    if ((e = tmp.get(&sort_tmp, sort_bytes)) != hipSuccess) return fail(e);
    if ((e = hipcub::DeviceRadixSort::SortPairs(sort_tmp, sort_bytes, keys, keys2, vals, order, (int)n, 0, 30,
                                                stream)) != hipSuccess)
        return fail(e);
    hipLaunchKernelGGL(k_leaves, dim3(blocks(n)), dim3(256), 0, stream, in.positions, in.indices, order, n, leaf_lo,
                       leaf_hi, out.tris, out.leaf_idx);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);

    if (m == 0) {   // one triangle
        if ((e = hipMalloc((void**)&out.nodes, 8u * sizeof(float4))) != hipSuccess) return fail(e);
        hipLaunchKernelGGL(k_single_root, dim3(1), dim3(1), 0, stream, leaf_lo, leaf_hi, out.nodes);
        if ((e = hipGetLastError()) != hipSuccess) return fail(e);
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) return fail(e);
        out.node_count = 1u;
        out.level_start = {0u, 1u};
        return hipSuccess;
    }

    // radix tree and its boxes
    uint32_t *left = nullptr, *right = nullptr, *arrivals = nullptr, *counters = nullptr;
    int32_t *parent_int = nullptr, *parent_leaf = nullptr;
    uint2 *range = nullptr, *qa = nullptr, *qb = nullptr;
    float* node_box = nullptr;
    float4* nodes = nullptr;
    if ((e = tmp.get(&left, m)) != hipSuccess || (e = tmp.get(&right, m)) != hipSuccess ||
        (e = tmp.get(&range, m)) != hipSuccess || (e = tmp.get(&parent_int, m)) != hipSuccess ||
        (e = tmp.get(&parent_leaf, n)) != hipSuccess || (e = tmp.get(&node_box, (size_t)m * 6u)) != hipSuccess ||
        (e = tmp.get(&arrivals, m)) != hipSuccess || (e = tmp.get(&counters, 2)) != hipSuccess ||
        (e = tmp.get(&qa, m)) != hipSuccess || (e = tmp.get(&qb, m)) != hipSuccess ||
        (e = tmp.get(&nodes, (size_t)m * 8u)) != hipSuccess)
        return fail(e);
    if ((e = hipMemsetAsync(parent_int, 0xFF, (size_t)m * sizeof(int32_t), stream)) != hipSuccess) return fail(e);
    if ((e = hipMemsetAsync(parent_leaf, 0xFF, (size_t)n * sizeof(int32_t), stream)) != hipSuccess) return fail(e);
    if ((e = hipMemsetAsync(arrivals, 0, (size_t)m * sizeof(uint32_t), stream)) != hipSuccess) return fail(e);
    hipLaunchKernelGGL(k_tree, dim3(blocks(m)), dim3(256), 0, stream, keys2, n, left, right, range, parent_int,
                       parent_leaf);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    hipLaunchKernelGGL(k_boxes, dim3(blocks(n)), dim3(256), 0, stream, n, left, right, parent_int, parent_leaf,
                       leaf_lo, leaf_hi, node_box, arrivals);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);

    // level-synchronous collapse from the root (radix-tree node 0 -> node 0)
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return fail(e);
    const uint2 root = make_uint2(0u, 0u);
    const uint32_t init[2] = {1u, 0u};
    if ((e = hipMemcpy(qa, &root, sizeof(root), hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
    if ((e = hipMemcpy(counters, init, sizeof(init), hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
    out.level_start = {0u, 1u};
    uint32_t count = 1u;
    while (count > 0u) {
        hipLaunchKernelGGL(k_collapse, dim3(blocks(count)), dim3(256), 0, stream, qa, count, left, right, range,
                           node_box, leaf_lo, leaf_hi, in.max_leaf, nodes, qb, counters);
        if ((e = hipGetLastError()) != hipSuccess) return fail(e);
        uint32_t c[2] = {0u, 0u};
        if ((e = hipMemcpyAsync(c, counters, sizeof(c), hipMemcpyDeviceToHost, stream)) != hipSuccess) return fail(e);
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) return fail(e);
        count = c[1];
        if (count > 0u) out.level_start.push_back(c[0]);
        if (c[0] > m) return fail(hipErrorUnknown);   // cannot happen: at most one node per radix-tree node
        if ((e = hipMemsetAsync(counters + 1, 0, sizeof(uint32_t), stream)) != hipSuccess) return fail(e);
        std::swap(qa, qb);
    }
    out.node_count = out.level_start.back();
    if ((e = hipMalloc((void**)&out.nodes, (size_t)out.node_count * 8u * sizeof(float4))) != hipSuccess) return fail(e);
    if ((e = hipMemcpyAsync(out.nodes, nodes, (size_t)out.node_count * 8u * sizeof(float4), hipMemcpyDeviceToDevice,
                            stream)) != hipSuccess)
        return fail(e);
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return fail(e);
    return hipSuccess;
}

}  // namespace dmt
