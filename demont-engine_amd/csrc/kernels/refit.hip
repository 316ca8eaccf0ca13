// BVH4 refit kernels: see refit.h.  Node4 = 8 float4: lo.x, hi.x, lo.y, hi.y,
// lo.z, hi.z rows of the four children, the child references, padding.
#include "refit.h"

#include <algorithm>

#include "dmt_dev.h"

namespace dmt {

namespace {

__device__ __forceinline__ f3 vertex(const float* pos, uint32_t i) {
    return f3{pos[3u * i], pos[3u * i + 1u], pos[3u * i + 2u]};
}

__global__ __launch_bounds__(256) void k_refit_tris(const float* __restrict__ pos, const uint32_t* __restrict__ idx,
                                                    uint32_t n, float4* __restrict__ tris) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const f3 a = vertex(pos, idx[3u * k]), b = vertex(pos, idx[3u * k + 1u]), c = vertex(pos, idx[3u * k + 2u]);
        float4* t = tris + 3u * (size_t)k;
        const float id = t[0].w;
        t[0] = make_float4(a.x, a.y, a.z, id);
        t[1] = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.0f);
        t[2] = make_float4(c.x - a.x, c.y - a.y, c.z - a.z, 0.0f);
    }
}

struct Box3 {
    f3 lo, hi;
};

__device__ __forceinline__ void grow(Box3& b, float lx, float hx, float ly, float hy, float lz, float hz) {
    b.lo = f3{fminf(b.lo.x, lx), fminf(b.lo.y, ly), fminf(b.lo.z, lz)};
    b.hi = f3{fmaxf(b.hi.x, hx), fmaxf(b.hi.y, hy), fmaxf(b.hi.z, hz)};
}

__device__ __forceinline__ float row(const float4& r, uint32_t i) {
    return i == 0u ? r.x : (i == 1u ? r.y : (i == 2u ? r.z : r.w));
}

__device__ Box3 child_box(const float4* nodes, uint32_t ref, const float* pos, const uint32_t* idx) {
    Box3 b{f3{INFINITY, INFINITY, INFINITY}, f3{-INFINITY, -INFINITY, -INFINITY}};
    if (ref & kLeaf) {
        const uint32_t first = ref & kLeafFirst, count = ((ref >> 27) & 15u) + 1u;
        for (uint32_t k = first; k < first + count; ++k) {
            const f3 v0 = vertex(pos, idx[3u * k]), v1 = vertex(pos, idx[3u * k + 1u]), v2 = vertex(pos, idx[3u * k + 2u]);
            // the host's item bounds: fmin(v0, fmin(v1, v2)), fmax likewise
            grow(b, fminf(v0.x, fminf(v1.x, v2.x)), fmaxf(v0.x, fmaxf(v1.x, v2.x)), fminf(v0.y, fminf(v1.y, v2.y)),
                 fmaxf(v0.y, fmaxf(v1.y, v2.y)), fminf(v0.z, fminf(v1.z, v2.z)), fmaxf(v0.z, fmaxf(v1.z, v2.z)));
        }
        return b;
    }
    const float4* n = nodes + 8u * (size_t)ref;
    const float4 lx = n[0], hx = n[1], ly = n[2], hy = n[3], lz = n[4], hz = n[5], ch = n[6];
#pragma unroll
    for (uint32_t i = 0; i < 4u; ++i) {
        if (__float_as_uint(row(ch, i)) == kEmpty) continue;
        grow(b, row(lx, i), row(hx, i), row(ly, i), row(hy, i), row(lz, i), row(hz, i));
    }
    return b;
}

__global__ __launch_bounds__(256) void k_refit_nodes(float4* __restrict__ nodes, const uint32_t* __restrict__ list,
                                                     uint32_t count, const float* __restrict__ pos,
                                                     const uint32_t* __restrict__ idx) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < count; j += stride) {
        float4* n = nodes + 8u * (size_t)list[j];
        const float4 ch = n[6];
        float r[6][4];
#pragma unroll
        for (uint32_t i = 0; i < 4u; ++i) {
            const uint32_t ref = __float_as_uint(row(ch, i));
            if (ref == kEmpty) {
                for (int a = 0; a < 6; ++a) r[a][i] = 0.0f;   // unused slots hold zero boxes, as built
                continue;
            }
            const Box3 b = child_box(nodes, ref, pos, idx);
            r[0][i] = b.lo.x; r[1][i] = b.hi.x;
            r[2][i] = b.lo.y; r[3][i] = b.hi.y;
            r[4][i] = b.lo.z; r[5][i] = b.hi.z;
        }
#pragma unroll
        for (int a = 0; a < 6; ++a) n[a] = make_float4(r[a][0], r[a][1], r[a][2], r[a][3]);
    }
}

uint32_t blocks_for(uint32_t n) { return std::max(1u, std::min((n + 255u) / 256u, 1u << 16)); }

}  // namespace

hipError_t launch_refit_tris(const float* positions, const uint32_t* leaf_idx, uint32_t tri_count, float4* tris,
                             hipStream_t stream) {
    if (tri_count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_refit_tris, dim3(blocks_for(tri_count)), dim3(256), 0, stream, positions, leaf_idx,
                       tri_count, tris);
    return hipGetLastError();
}

hipError_t launch_refit_nodes(float4* nodes, const uint32_t* node_list, uint32_t count, const float* positions,
                              const uint32_t* leaf_idx, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_refit_nodes, dim3(blocks_for(count)), dim3(256), 0, stream, nodes, node_list, count,
                       positions, leaf_idx);
    return hipGetLastError();
}

}  // namespace dmt
