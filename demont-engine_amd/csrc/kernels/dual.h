// Dual numbers for the adjoint's tangent pass (forward mode): df{v, d}
// carries a value and its derivative along one parameter direction.
//
// Every operation computes its value with exactly the float operation the
// rendering code uses (same operands, same order, the same fmaxf / Cephes
// routines), so a dual run takes the float run's decisions and its values
// equal the renderer's bit for bit; comparisons look at the value only.
#pragma once
#include "dmt_dev.h"

namespace dmt {

struct df {
    float v, d;
    df() = default;
    __host__ __device__ constexpr df(float x) : v(x), d(0.0f) {}
    __host__ __device__ constexpr df(float x, float dx) : v(x), d(dx) {}
};

__device__ __forceinline__ df operator+(df a, df b) { return df(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ df operator-(df a, df b) { return df(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ df operator-(df a) { return df(-a.v, -a.d); }
__device__ __forceinline__ df operator*(df a, df b) { return df(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ df operator/(df a, df b) {
    const float q = a.v / b.v;
    return df(q, (a.d - q * b.d) / b.v);
}
__device__ __forceinline__ bool operator<(df a, df b) { return a.v < b.v; }
__device__ __forceinline__ bool operator>(df a, df b) { return a.v > b.v; }
__device__ __forceinline__ bool operator<=(df a, df b) { return a.v <= b.v; }
__device__ __forceinline__ bool operator>=(df a, df b) { return a.v >= b.v; }

__device__ __forceinline__ df sqrt_(df a) {
    const float s = sqrtf(a.v);
    return df(s, s > 0.0f ? a.d / (2.0f * s) : 0.0f);
}
__device__ __forceinline__ df fmax_(df a, df b) {
    const float m = fmaxf(a.v, b.v);
    return df(m, m == a.v ? a.d : b.d);
}
__device__ __forceinline__ df fmin_(df a, df b) {
    const float m = fminf(a.v, b.v);
    return df(m, m == a.v ? a.d : b.d);
}
__device__ __forceinline__ df fabs_(df a) { return df(fabsf(a.v), a.v < 0.0f ? -a.d : a.d); }
__device__ __forceinline__ float primal(df a) { return a.v; }

__device__ __forceinline__ df c_atan2(df y, df x) {
    const float r2 = x.v * x.v + y.v * y.v;
    return df(c_atan2(y.v, x.v), r2 > 0.0f ? (x.v * y.d - y.v * x.d) / r2 : 0.0f);
}
__device__ __forceinline__ df c_acos(df x) {
    const float s = sqrtf(fmaxf(1.0f - x.v * x.v, 0.0f));
    return df(c_acos(x.v), s > 0.0f ? -x.d / s : 0.0f);
}

using d3 = V3<df>;

__device__ __forceinline__ f3 value3(d3 a) { return f3{a.x.v, a.y.v, a.z.v}; }
__device__ __forceinline__ f3 deriv3(d3 a) { return f3{a.x.d, a.y.d, a.z.d}; }

// Surface of a hit along a dual ray: the hit distance keeps the traversal's
// value h.t and takes its derivative from the primitive's implicit equation
// g(o + t d) = 0, dt = -(dg . do + t dg . dd) / (dg . d), with dg the sphere's
// radius vector or the plane / triangle normal at the hit.  A sphere is taken
// about its centre at shutter time `time` (PathTrace.slang:2093).  A hit on an
// SDF cluster (SDF) is held fixed: the clusters carry no differentiated
// parameter and their hit distance and normal carry no tangent.
template <bool SDF = false>
__device__ __forceinline__ SurfaceT<df> surface_of(const SceneView& S, d3 o, d3 d, HitRec h, float time = 1.0f) {
    SurfaceT<df> s;
    surface_material<SDF>(S, h, s);
    const f3 ov = value3(o), dv = value3(d), od = deriv3(o), dd = deriv3(d);
    f3 g;
    bool sphere = false;
    f3 centre{0.0f, 0.0f, 0.0f};
    if (h.code & kMeshBit) {
        const float4* tr = S.tris + (size_t)(h.code & ~kMeshBit) * 3u;
        g = cross(xyz(tr[1]), xyz(tr[2]));
        s.n = to<df>(normalize(g));
    } else if (SDF && (h.code & kSdfBit)) {
        const float4* c = S.sdf + (size_t)(h.code & ~kSdfBit) * kSdfStride;
        s.t = df(h.t, 0.0f);
        s.n = to<df>(normalize(sdf_grad(c, __float_as_uint(c[0].w), ov + dv * h.t)));
        return s;
    } else {   // analytic
        const float4* p = S.prims + (size_t)h.code * 7u;
        const float4 v0 = p[0], v2 = p[2];
        if (v2.x == 0.0f) {
            sphere = true;
            centre = xyz(v0);
            if (time < 1.0f) centre = mix3(xyz(p[3]), centre, time);
            g = (ov + dv * h.t) - centre;
        } else {
            g = quat_rotate(p[5], xyz(v0));
            s.n = to<df>(g);
        }
    }
    const float den = dot(g, dv);
    const float dt = den != 0.0f ? -(dot(g, od) + h.t * dot(g, dd)) / den : 0.0f;
    s.t = df(h.t, dt);
    if (sphere) s.n = normalize((o + d * s.t) - to<df>(centre));
    return s;
}

}  // namespace dmt
