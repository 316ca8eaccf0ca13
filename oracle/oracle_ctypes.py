"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as the
checker / baseline.  The product never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


class OrcSceneDesc(C.Structure):
    _fields_ = [
        ("prims", C.POINTER(C.c_float)), ("prim_count", C.c_uint32),
        ("lights", C.POINTER(C.c_float)), ("light_count", C.c_uint32),
        ("positions", C.POINTER(C.c_float)), ("vertex_count", C.c_uint32),
        ("indices", C.POINTER(C.c_uint32)), ("index_count", C.c_uint32),
        ("env_rgb", C.POINTER(C.c_float)), ("env_width", C.c_uint32), ("env_height", C.c_uint32),
        ("env_intensity", C.c_float),
        ("flags", C.c_uint32),
        ("env_marginal", C.POINTER(C.c_float)), ("env_conditional", C.POINTER(C.c_float)),
        ("env_total_luminance", C.c_double),
        ("moon_rgba", C.POINTER(C.c_float)), ("moon_width", C.c_uint32), ("moon_height", C.c_uint32),
        ("sdf_clusters", C.POINTER(C.c_float)), ("sdf_cluster_count", C.c_uint32),
        ("pbr_atlas", C.POINTER(C.c_uint8)), ("pbr_atlas_width", C.c_uint32), ("pbr_atlas_height", C.c_uint32),
        ("mesh_uvs", C.POINTER(C.c_float)),
    ]


class OrcCounts(C.Structure):
    _fields_ = [("primary", C.c_uint64), ("extension", C.c_uint64), ("shadow", C.c_uint64),
                ("node_visits", C.c_uint64), ("prim_tests", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


HIT_DTYPE = np.dtype([("t", "<f4"), ("normal", "<f4", (3,)), ("hit", "<u4"), ("prim", "<u4"),
                      ("mat", "<u4"), ("user_id", "<u4")])

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, u32, fp = C.c_void_p, C.c_uint32, C.POINTER(C.c_float)
        L.orc_scene_create.argtypes = [C.POINTER(OrcSceneDesc), C.POINTER(vp)]
        L.orc_scene_create.restype = C.c_int
        L.orc_scene_destroy.argtypes = [vp]
        L.orc_scene_destroy.restype = None
        L.orc_render.argtypes = [vp, vp, C.c_int, C.c_int, u32, u32, u32, u32, fp, C.POINTER(OrcCounts)]
        L.orc_render.restype = C.c_int
        L.orc_trace.argtypes = [vp, fp, u32, vp]
        L.orc_trace.restype = C.c_int
        L.orc_trace_bruteforce.argtypes = [vp, fp, u32, vp]
        L.orc_gbuffer.restype = C.c_int
        L.orc_gbuffer.argtypes = [vp, vp, fp, fp, fp]
        L.orc_trace_bruteforce.restype = C.c_int
        L.orc_build_env_cdf.argtypes = [fp, u32, u32, fp, fp]
        L.orc_build_env_cdf.restype = C.c_double
        L.orc_build_light_tree.argtypes = [fp, u32, fp]
        L.orc_build_light_tree.restype = u32
        L.orc_build_tri_bvh.argtypes = [fp, u32, C.POINTER(u32), u32, C.POINTER(u32), C.POINTER(u32)]
        L.orc_build_tri_bvh.restype = u32
        L.orc_env_sample.argtypes = [vp, fp, u32, fp]
        L.orc_env_sample.restype = C.c_int
        L.orc_env_pdf.argtypes = [vp, fp, u32, fp]
        L.orc_env_pdf.restype = C.c_int
        L.orc_light_pmf.argtypes = [vp, fp, fp, fp]
        L.orc_light_pmf.restype = C.c_int
        L.orc_math_f32.argtypes = [C.c_int, fp, fp, fp, u32]
        L.orc_math_f32.restype = None
        L.orc_tangent.argtypes = [vp, vp, u32, u32, fp, C.c_int, C.POINTER(C.c_double), fp]
        L.orc_tangent.restype = C.c_int
        L.orc_render_f64_image.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_double)]
        L.orc_render_f64_image.restype = C.c_int
        L.orc_trace_at.argtypes = [vp, fp, u32, C.c_float, fp, fp, C.c_int, vp]
        L.orc_trace_at.restype = C.c_int
        L.orc_lens_sample.argtypes = [fp, fp, u32, C.c_int, C.c_float, C.c_float, fp]
        L.orc_lens_sample.restype = C.c_int
        L.orc_sky.argtypes = [vp, vp, fp, u32, fp]
        L.orc_sky.restype = C.c_int
        L.orc_sample_sun_disc.argtypes = [fp, u32, u32, fp]
        L.orc_sample_sun_disc.restype = None
        L.orc_sdf_eval.argtypes = [fp, fp, u32, fp]
        L.orc_sdf_eval.restype = None
        L.orc_sdf_trace.argtypes = [fp, fp, u32, u32, C.c_float, fp]
        L.orc_sdf_trace.restype = None
        L.orc_pbr_sample.argtypes = [vp, u32, fp, u32, fp]
        L.orc_pbr_sample.restype = C.c_int
        L.orc_pbr_surface.argtypes = [vp, vp, fp, u32, C.c_float, fp]
        L.orc_pbr_surface.restype = C.c_int
        L.orc_pbr_normal_map.argtypes = [fp, fp, u32, fp]
        L.orc_pbr_normal_map.restype = None
        L.orc_pbr_srgb.argtypes = [fp, u32, fp]
        L.orc_pbr_srgb.restype = None
        L.orc_value_noise.argtypes = [fp, u32, fp]
        L.orc_value_noise.restype = None
        L.orc_water_normal.argtypes = [vp, fp, fp, u32, fp]
        L.orc_water_normal.restype = C.c_int
        _lib = L
    return _lib


def _fp(a):
    if a is None or a.size == 0:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_float))


class OracleScene:
    """Oracle copy of a scene given in the reference wire formats."""

    def __init__(self, prims, lights, positions, indices, env=None, env_intensity=1.0, env_cdf=None, moon=None,
                 sdf=None, atlas=None, uvs=None):
        self._keep = [np.ascontiguousarray(prims, np.float32), np.ascontiguousarray(lights, np.float32),
                      np.ascontiguousarray(positions, np.float32), np.ascontiguousarray(indices, np.uint32)]
        d = OrcSceneDesc()
        p, l, v, i = self._keep
        d.prims, d.prim_count = _fp(p), p.shape[0] if p.ndim > 1 else p.size // 28
        d.lights, d.light_count = _fp(l), l.shape[0] if l.ndim > 1 else l.size // 20
        d.positions, d.vertex_count = _fp(v), v.size // 3
        d.indices = i.ctypes.data_as(C.POINTER(C.c_uint32)) if i.size else None
        d.index_count = i.size
        if env is not None:
            e = np.ascontiguousarray(env, np.float32)
            self._keep.append(e)
            d.env_rgb, d.env_height, d.env_width = _fp(e), e.shape[0], e.shape[1]
        d.env_intensity = env_intensity
        if env_cdf is not None:
            m = np.ascontiguousarray(env_cdf[0], np.float32)
            c = np.ascontiguousarray(env_cdf[1], np.float32)
            self._keep += [m, c]
            d.env_marginal, d.env_conditional, d.env_total_luminance = _fp(m), _fp(c), float(env_cdf[2])
        if moon is not None:
            mo = np.ascontiguousarray(moon, np.float32)
            self._keep.append(mo)
            d.moon_rgba, d.moon_height, d.moon_width = _fp(mo), mo.shape[0], mo.shape[1]
        if sdf is not None and np.asarray(sdf).size:
            sc = np.ascontiguousarray(sdf, np.float32).reshape(-1, 112)
            self._keep.append(sc)
            d.sdf_clusters, d.sdf_cluster_count = _fp(sc), sc.shape[0]
        if atlas is not None:
            at = np.ascontiguousarray(atlas, np.uint8)
            self._keep.append(at)
            d.pbr_atlas = at.ctypes.data_as(C.POINTER(C.c_uint8))
            d.pbr_atlas_height, d.pbr_atlas_width = at.shape[0], at.shape[1]
        if uvs is not None:
            uv = np.ascontiguousarray(uvs, np.float32).reshape(-1, 2)
            self._keep.append(uv)
            d.mesh_uvs = _fp(uv)
        self._h = C.c_void_p()
        rc = lib().orc_scene_create(C.byref(d), C.byref(self._h))
        if rc != 0:
            raise ValueError("oracle rejected the scene")

    @classmethod
    def from_scene_data(cls, sd):
        return cls(sd.prims, sd.lights, sd.positions, sd.indices, sd.env, sd.env_intensity,
                   getattr(sd, "env_cdf", None), getattr(sd, "moon", None), getattr(sd, "sdf", None),
                   getattr(sd, "atlas", None), getattr(sd, "uvs", None))

    def render(self, frame, precision=32, threads=0, rows=None, col_step=1, accum=None):
        """Render one frame; `frame` is any ctypes struct with the dmt_frame_params layout.
        rows = (begin, end, step).  Returns (accum (H, W, 4), counts dict)."""
        W, H = frame.width, frame.height
        if threads <= 0:
            threads = min(16, os.cpu_count() or 1)
        if accum is None:
            accum = np.zeros((H, W, 4), np.float32)
        r0, r1, rs = rows if rows is not None else (0, H, 1)
        cnt = OrcCounts()
        rc = lib().orc_render(self._h, C.cast(C.byref(frame), C.c_void_p), precision, threads, r0, r1, rs, col_step,
                              _fp(accum), C.byref(cnt))
        if rc != 0:
            raise ValueError("oracle render failed")
        return accum, cnt.as_dict()

    def tangent(self, frame, prim, param, dimage, threads=0, want_image=False):
        """fp64 tangent-linear derivative sum_p dimage[p] . dI[p]/d(param of prim)
        (param 0-2 albedo rgb, 3 roughness, 4-6 emission rgb)."""
        if threads <= 0:
            threads = min(16, os.cpu_count() or 1)
        di = np.ascontiguousarray(dimage, np.float32)
        out = C.c_double()
        img = np.zeros((frame.height, frame.width, 4), np.float32) if want_image else None
        rc = lib().orc_tangent(self._h, C.cast(C.byref(frame), C.c_void_p), prim, param, _fp(di), threads,
                               C.byref(out), _fp(img) if want_image else None)
        if rc != 0:
            raise ValueError("oracle tangent failed")
        return (out.value, img) if want_image else out.value

    def render_f64(self, frame, threads=0):
        """fp64 frame value per pixel (H, W, 3) doubles."""
        if threads <= 0:
            threads = min(16, os.cpu_count() or 1)
        out = np.zeros((frame.height, frame.width, 3), np.float64)
        rc = lib().orc_render_f64_image(self._h, C.cast(C.byref(frame), C.c_void_p), threads,
                                        out.ctypes.data_as(C.POINTER(C.c_double)))
        if rc != 0:
            raise ValueError("oracle f64 render failed")
        return out

    def trace(self, origins, dirs, brute=False):
        n = origins.shape[0]
        rays = np.zeros((n, 8), np.float32)
        rays[:, 0:3] = origins
        rays[:, 4:7] = dirs
        out = np.zeros(n, HIT_DTYPE)
        fn = lib().orc_trace_bruteforce if brute else lib().orc_trace
        fn(self._h, _fp(rays), n, out.ctypes.data_as(C.c_void_p))
        return out

    def trace_at(self, origins, dirs, time, shift=None, brute=False):
        """Closest hits at shutter time `time`; shift = (prev xyz, curr xyz) of
        a mesh translation or None."""
        n = origins.shape[0]
        rays = np.zeros((n, 8), np.float32)
        rays[:, 0:3] = origins
        rays[:, 4:7] = dirs
        out = np.zeros(n, HIT_DTYPE)
        sp = sc = None
        if shift is not None:
            sp = np.ascontiguousarray(shift[0], np.float32)
            sc = np.ascontiguousarray(shift[1], np.float32)
        rc = lib().orc_trace_at(self._h, _fp(rays), n, float(time), _fp(sp), _fp(sc), 1 if brute else 0,
                                out.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise ValueError("oracle trace failed")
        return out

    def gbuffer(self, frame, vp_curr, vp_prev):
        """(H, W, 16) float32: depth, motion.xy, normal.xyz, albedo.xyz, F0.xyz,
        roughness, specular hit distance, 0, 0 (column-major view-projections)."""
        out = np.zeros((frame.height, frame.width, 16), np.float32)
        a = np.ascontiguousarray(vp_curr, np.float32).reshape(16)
        b = np.ascontiguousarray(vp_prev, np.float32).reshape(16)
        rc = lib().orc_gbuffer(self._h, C.cast(C.byref(frame), C.c_void_p), _fp(a), _fp(b), _fp(out))
        if rc != 0:
            raise ValueError("oracle gbuffer failed")
        return out

    def sky(self, frame, dirs):
        """skyColor of missed rays along unit dirs (n, 3) under the frame's sky lights."""
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        out = np.zeros_like(d)
        if lib().orc_sky(self._h, C.cast(C.byref(frame), C.c_void_p), _fp(d), d.shape[0], _fp(out)) != 0:
            raise ValueError("oracle sky failed")
        return out

    def env_sample(self, xi):
        xi = np.ascontiguousarray(xi, np.float32)
        n = xi.shape[0]
        out = np.zeros((n, 8), np.float32)
        lib().orc_env_sample(self._h, _fp(xi), n, _fp(out))
        return out[:, 0:3], out[:, 3:6], out[:, 6]

    def env_pdf(self, dirs):
        d = np.ascontiguousarray(dirs, np.float32)
        out = np.zeros(d.shape[0], np.float32)
        lib().orc_env_pdf(self._h, _fp(d), d.shape[0], _fp(out))
        return out

    def light_pmf(self, p, nf, n_lights):
        pv = np.ascontiguousarray(p, np.float32)
        nv = np.ascontiguousarray(nf, np.float32)
        out = np.zeros(n_lights, np.float32)
        if lib().orc_light_pmf(self._h, _fp(pv), _fp(nv), _fp(out)) != 0:
            raise ValueError("no light tree")
        return out

    def pbr_sample(self, tile, uvl):
        """samplePbrTileLinear of `tile` at (u, v, lod) rows (n, 3): (n, 4)."""
        q = np.ascontiguousarray(uvl, np.float32).reshape(-1, 3)
        out = np.zeros((q.shape[0], 4), np.float32)
        if lib().orc_pbr_sample(self._h, tile, _fp(q), q.shape[0], _fp(out)) != 0:
            raise ValueError("no atlas or tile out of range")
        return out

    def pbr_surface(self, frame, origins, dirs, cone_width=0.0):
        """The hit's PBR surface (n, 16): u, v, uv_scale, lod, albedo rgb,
        roughness, material, normal xyz, t, kind, 0, 0."""
        n = np.asarray(origins).reshape(-1, 3).shape[0]
        rays = np.zeros((n, 8), np.float32)
        rays[:, 0:3] = np.asarray(origins, np.float32).reshape(-1, 3)
        rays[:, 4:7] = np.asarray(dirs, np.float32).reshape(-1, 3)
        out = np.zeros((n, 16), np.float32)
        if lib().orc_pbr_surface(self._h, C.cast(C.byref(frame), C.c_void_p), _fp(rays), n, float(cone_width),
                                 _fp(out)) != 0:
            raise ValueError("oracle pbr surface failed")
        return out

    def __del__(self):
        try:
            if self._h:
                lib().orc_scene_destroy(self._h)
        except Exception:
            pass


def env_cdf(env):
    e = np.ascontiguousarray(env, np.float32)
    H, W = e.shape[:2]
    m = np.zeros(H, np.float32)
    c = np.zeros(H * W, np.float32)
    t = lib().orc_build_env_cdf(_fp(e), W, H, _fp(m), _fp(c))
    return m, c.reshape(H, W), t


def light_tree(lights):
    L = np.ascontiguousarray(lights, np.float32)
    n = L.shape[0]
    out = np.zeros((max(2 * n - 1, 1), 16), np.float32)
    k = lib().orc_build_light_tree(_fp(L), n, _fp(out))
    return out[:k]


def tri_bvh(positions, indices):
    p = np.ascontiguousarray(positions, np.float32)
    i = np.ascontiguousarray(indices, np.uint32).reshape(-1)
    t = i.size // 3
    nodes = np.zeros((max(2 * t, 1), 8), np.uint32)
    perm = np.zeros(max(t, 1), np.uint32)
    k = lib().orc_build_tri_bvh(_fp(p), p.size // 3, i.ctypes.data_as(C.POINTER(C.c_uint32)), i.size,
                                nodes.ctypes.data_as(C.POINTER(C.c_uint32)), perm.ctypes.data_as(C.POINTER(C.c_uint32)))
    return nodes[:k], perm


def math_f32(fn, a, b=None):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b if b is not None else np.zeros_like(a), np.float32)
    out = np.zeros_like(a)
    code = {"sin": 0, "cos": 1, "acos": 2, "atan2": 3, "exp": 4, "spectral_r": 5, "spectral_g": 6,
            "spectral_b": 7, "cauchy": 8, "log": 9, "pow": 10}[fn]
    lib().orc_math_f32(code, _fp(a), _fp(b), _fp(out), a.size)
    return out


def lens_sample(u, blades=0, rot=0.0, cat_eye=0.0, uv=None):
    """Lens points (n, 2) of the uniforms u (n, 2): the concentric disk
    (blades < 3) or the rotated regular polygon, then the cat's-eye squeeze
    at screen positions uv (n, 2) when cat_eye > 0."""
    u = np.ascontiguousarray(u, np.float32).reshape(-1, 2)
    uvv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2) if uv is not None else None
    out = np.zeros_like(u)
    rc = lib().orc_lens_sample(_fp(u), _fp(uvv) if uvv is not None else None, u.shape[0], int(blades), float(rot),
                               float(cat_eye), _fp(out))
    if rc != 0:
        raise ValueError("lens sample rejected")
    return out


def sample_sun_disc(direction, seed0, n):
    """sampleSunDisc around the unit direction with seeds seed0 .. seed0 + n - 1."""
    d = np.ascontiguousarray(direction, np.float32)
    out = np.zeros((n, 3), np.float32)
    lib().orc_sample_sun_disc(_fp(d), seed0, n, _fp(out))
    return out


def sdf_eval(cluster, points):
    """(distance, unit normal) of one 112-float SDF cluster at points (n, 3): (n, 4)."""
    c = np.ascontiguousarray(cluster, np.float32).reshape(112)
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    out = np.zeros((p.shape[0], 4), np.float32)
    lib().orc_sdf_eval(_fp(c), _fp(p), p.shape[0], _fp(out))
    return out


def sdf_trace(cluster, origins, dirs, max_iters=128, epsilon=1e-4):
    """sdfClusterTrace of one cluster (t_min 1e-3, t_max 1e30): (n, 4) of
    (t or -1, normal xyz)."""
    c = np.ascontiguousarray(cluster, np.float32).reshape(112)
    n = np.asarray(origins).reshape(-1, 3).shape[0]
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = np.asarray(origins, np.float32).reshape(-1, 3)
    rays[:, 4:7] = np.asarray(dirs, np.float32).reshape(-1, 3)
    out = np.zeros((n, 4), np.float32)
    lib().orc_sdf_trace(_fp(c), _fp(rays), n, max_iters, epsilon, _fp(out))
    return out


def pbr_normal_map(normals, ts):
    """applyNormalMap (fp32) of unit normals (n, 3) by tangent-space vectors (n, 3)."""
    n = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(ts, np.float32).reshape(-1, 3)
    out = np.zeros_like(n)
    lib().orc_pbr_normal_map(_fp(n), _fp(t), n.shape[0], _fp(out))
    return out


def pbr_srgb(c):
    """srgbToLinear (fp32) of values c."""
    c = np.ascontiguousarray(c, np.float32).reshape(-1)
    out = np.zeros_like(c)
    lib().orc_pbr_srgb(_fp(c), c.size, _fp(out))
    return out


def value_noise(points):
    """cloud_value_noise (fp32) at points (n, 3)."""
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    out = np.zeros(p.shape[0], np.float32)
    lib().orc_value_noise(_fp(p), p.shape[0], _fp(out))
    return out


def water_normal(frame, points, normals):
    """waterShadingNormal (fp32) of the frame's wave parameters at points (n, 3) with macro normals (n, 3)."""
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    n = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    out = np.zeros_like(p)
    if lib().orc_water_normal(C.cast(C.byref(frame), C.c_void_p), _fp(p), _fp(n), p.shape[0], _fp(out)) != 0:
        raise ValueError("oracle water normal failed")
    return out
