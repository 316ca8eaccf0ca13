"""ctypes binding of libdmt_hip.so (the C ABI declared in include/dmt.h).

Thin plumbing for the test harness and the benchmark: every call goes through
the C ABI of the in-tree library.  Loading fails loudly when the library has
not been built -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "libdmt_hip.so"))

DMT_OK = 0
DMT_ERR_INVALID_ARG = -1
DMT_ERR_OUT_OF_MEMORY = -2
DMT_ERR_NO_DEVICE = -3
DMT_ERR_DEVICE = -4
DMT_ERR_UNSUPPORTED = -5
DMT_ERR_DEVICE_LOST = -6
DMT_ERR_LIMIT = -7

ABI_VERSION = 8
DMT_FRAME_NO_RUSSIAN_ROULETTE = 1
DMT_FRAME_FIXED_JITTER = 2
FRAME_MESH_TEXTURES = 4          # DMT_FRAME_MESH_TEXTURES
MAT_WATER = 3                    # DMT_MAT_WATER
SCHEDULE_WAVEFRONT = 0xFFFFFFFF
MATERIAL_FLOATS = 7
DMT_NOALLOC_BLOCKS = 512        # dmt_trace_closest_noalloc's grid
DMT_NOALLOC_SPILL_DEPTH = 48    # and its static per-thread stack spill

PRIM_FLOATS = 28
LIGHT_FLOATS = 20
SDF_CLUSTER_FLOATS = 112


class FrameParams(C.Structure):
    _fields_ = [
        ("pos_fovtan", C.c_float * 4),
        ("fwd_aspect", C.c_float * 4),
        ("right_xyz", C.c_float * 4),
        ("up_xyz", C.c_float * 4),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("frame_index", C.c_uint32),
        ("spp", C.c_uint32),
        ("max_bounces", C.c_uint32),
        ("reset_accum", C.c_uint32),
        ("mis_enabled", C.c_uint32),
        ("light_tree_enabled", C.c_uint32),
        ("refract_bounces", C.c_uint32),
        ("firefly_clamp", C.c_float),
        ("ema_alpha", C.c_float),
        ("flags", C.c_uint32),
        ("dof", C.c_float * 4),
        ("cat_eye", C.c_float),
        ("motion_blur", C.c_uint32),
        ("spectral", C.c_uint32),
        ("cauchy_c", C.c_float),
        ("mesh_shift_prev", C.c_float * 3),
        ("mesh_shift_active", C.c_uint32),
        ("mesh_shift_curr", C.c_float * 3),
        ("reserved", C.c_uint32),
        ("jitter", C.c_float * 2),
        ("sun", C.c_float * 4),
        ("sun_extra", C.c_float * 4),
        ("sun_disc_brightness", C.c_float),
        ("moon", C.c_float * 4),
        ("moon_extra", C.c_float * 4),
        ("hdri_light_count", C.c_uint32),
        ("hdri_dir", (C.c_float * 4) * 8),
        ("hdri_col", (C.c_float * 4) * 8),
        ("sdf_max_iters", C.c_uint32),
        ("sdf_epsilon", C.c_float),
        ("mesh_tex", C.c_uint32 * 4),
        ("water_absorb", C.c_float * 3),
        ("water_ior", C.c_float),
        ("wave", C.c_float * 4),
        ("camera_in_water", C.c_uint32),
        ("deep_color", C.c_float * 3),
    ]

    def copy(self) -> "FrameParams":
        f = FrameParams()
        C.memmove(C.byref(f), C.byref(self), C.sizeof(FrameParams))
        return f


class SceneDesc(C.Structure):
    _fields_ = [
        ("prims", C.POINTER(C.c_float)), ("prim_count", C.c_uint32),
        ("lights", C.POINTER(C.c_float)), ("light_count", C.c_uint32),
        ("positions", C.POINTER(C.c_float)), ("vertex_count", C.c_uint32),
        ("indices", C.POINTER(C.c_uint32)), ("index_count", C.c_uint32),
        ("env_rgb", C.POINTER(C.c_float)), ("env_width", C.c_uint32), ("env_height", C.c_uint32),
        ("env_intensity", C.c_float),
        ("reserved", C.c_uint32),
        ("env_marginal", C.POINTER(C.c_float)),
        ("env_conditional", C.POINTER(C.c_float)),
        ("env_total_luminance", C.c_double),
        ("sdf_clusters", C.POINTER(C.c_float)), ("sdf_cluster_count", C.c_uint32),
        ("reserved2", C.c_uint32),
        ("mesh_uvs", C.POINTER(C.c_float)),
    ]


class BuildParams(C.Structure):
    _fields_ = [("sah_bins", C.c_uint32), ("max_leaf", C.c_uint32), ("flags", C.c_uint32), ("reserved", C.c_uint32)]


BUILD_DEVICE = 1


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("normal", C.c_float * 3), ("hit", C.c_uint32), ("prim", C.c_uint32),
                ("mat", C.c_uint32), ("user_id", C.c_uint32)]


HIT_DTYPE = np.dtype([("t", "<f4"), ("normal", "<f4", (3,)), ("hit", "<u4"), ("prim", "<u4"),
                      ("mat", "<u4"), ("user_id", "<u4")])


class Stats(C.Structure):
    _fields_ = [
        ("primary_rays", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
        ("frames", C.c_uint64), ("extend_launches", C.c_uint64), ("extend_ms", C.c_double),
        ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64), ("sphere_tests", C.c_uint64),
        ("ext_node_visits", C.c_uint64), ("ext_tri_tests", C.c_uint64), ("ext_sphere_tests", C.c_uint64),
        ("scene_bytes", C.c_uint64), ("bvh_nodes", C.c_uint32), ("bvh_depth", C.c_uint32),
        ("tail_node_visits", C.c_uint64), ("tail_tri_tests", C.c_uint64), ("tail_sphere_tests", C.c_uint64),
        ("ext_rays", C.c_uint64), ("tail_launches", C.c_uint64), ("tail_ms", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class GenOptions(C.Structure):
    _fields_ = [("subdivisions", C.c_int32), ("random_tris", C.c_int32), ("spheres", C.c_int32),
                ("width", C.c_uint32), ("height", C.c_uint32), ("env_width", C.c_uint32),
                ("env_height", C.c_uint32), ("lights", C.c_int32), ("sdf_clusters", C.c_int32)]


class GeneratedScene(C.Structure):
    _fields_ = [
        ("prims", C.POINTER(C.c_float)), ("prim_count", C.c_uint32),
        ("lights", C.POINTER(C.c_float)), ("light_count", C.c_uint32),
        ("positions", C.POINTER(C.c_float)), ("vertex_count", C.c_uint32),
        ("indices", C.POINTER(C.c_uint32)), ("index_count", C.c_uint32),
        ("env_rgb", C.POINTER(C.c_float)), ("env_width", C.c_uint32), ("env_height", C.c_uint32),
        ("env_intensity", C.c_float),
        ("frame", FrameParams),
        ("sdf_clusters", C.POINTER(C.c_float)), ("sdf_cluster_count", C.c_uint32),
    ]


class BatchParams(C.Structure):
    _fields_ = [("tile_w", C.c_uint32), ("tiles_per_fetch", C.c_uint32), ("blocks_per_cu", C.c_uint32),
                ("stack_lds", C.c_uint32), ("plain_launches", C.c_uint32), ("reserved", C.c_uint32 * 3)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class TonemapParams(C.Structure):
    _fields_ = [("exposure", C.c_float), ("op", C.c_uint32), ("mode", C.c_uint32), ("reserved", C.c_uint32)]


TONEMAP_DISPLAY, TONEMAP_LINEAR, TONEMAP_HDR = 0, 1, 2


class GbufferParams(C.Structure):
    _fields_ = [("curr_view_proj", C.c_float * 16), ("prev_view_proj", C.c_float * 16),
                ("pixel_stride", C.c_uint32), ("reserved", C.c_uint32),
                ("depth", C.c_void_p), ("motion", C.c_void_p), ("normal", C.c_void_p), ("albedo", C.c_void_p),
                ("specular_albedo", C.c_void_p), ("roughness", C.c_void_p), ("specular_hit_distance", C.c_void_p),
                ("cloud_transmittance", C.c_void_p)]


class DmtError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"dmt status {status}: {msg}")
        self.status = status


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
    # PyTorch-ROCm ships its own HIP runtime.  Two runtimes in one process
    # cannot both open the GPU: the one initialised second reports no device.
    # Loading torch first makes this library bind to torch's runtime, so
    # device pointers pass between the two.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    v = C.c_uint32()
    if L.dmt_get_abi_version(C.byref(v)) != 0 or v.value != ABI_VERSION:
        raise ImportError(f"{LIB_PATH} has ABI version {v.value}, this binding expects {ABI_VERSION}: rebuild it")
    vp, u32, i32, st = C.c_void_p, C.c_uint32, C.c_int32, C.c_int32
    fp = C.POINTER(C.c_float)
    sigs = {
        "dmt_get_abi_version": (st, [C.POINTER(u32)]),
        "dmt_device_create": (st, [i32, C.POINTER(vp)]),
        "dmt_device_destroy": (None, [vp]),
        "dmt_device_stream": (st, [vp, C.POINTER(vp)]),
        "dmt_device_synchronize": (st, [vp]),
        "dmt_device_is_lost": (i32, [vp]),
        "dmt_scene_create": (st, [vp, C.POINTER(SceneDesc), C.POINTER(BuildParams), C.POINTER(vp)]),
        "dmt_scene_destroy": (None, [vp]),
        "dmt_render": (st, [vp, C.POINTER(FrameParams)]),
        "dmt_render_frames": (st, [vp, C.POINTER(FrameParams), u32]),
        "dmt_readback_accum": (st, [vp, fp, C.c_size_t]),
        "dmt_accum_device_ptr": (st, [vp, C.POINTER(vp)]),
        "dmt_scene_bind_accum": (st, [vp, vp, u32, u32]),
        "dmt_tonemap": (st, [vp, C.POINTER(TonemapParams), vp]),
        "dmt_render_gbuffer": (st, [vp, C.POINTER(FrameParams), C.POINTER(GbufferParams)]),
        "dmt_scene_update_mesh": (st, [vp, fp, u32]),
        "dmt_trace_closest": (st, [vp, fp, u32, vp]),
        "dmt_trace_closest_noalloc": (st, [vp, vp, u32, vp]),
        "dmt_scene_stats": (st, [vp, C.POINTER(Stats)]),
        "dmt_scene_reset_stats": (st, [vp]),
        "dmt_scene_measure_traffic": (st, [vp, C.POINTER(FrameParams), C.POINTER(Stats)]),
        "dmt_scene_set_profiling": (st, [vp, i32]),
        "dmt_scene_set_schedule": (st, [vp, u32]),
        "dmt_scene_get_schedule": (st, [vp, C.POINTER(u32)]),
        "dmt_last_error": (C.c_char_p, []),
        "dmt_scene_generate": (st, [u32, C.c_uint64, C.POINTER(GenOptions), C.POINTER(C.POINTER(GeneratedScene))]),
        "dmt_scene_generated_free": (None, [C.POINTER(GeneratedScene)]),
        "dmt_camera_look_at": (st, [fp, fp, C.c_float, u32, u32, C.POINTER(FrameParams)]),
        "dmt_build_bvh4": (st, [fp, u32, C.POINTER(u32), u32, u32, u32, vp, vp, C.POINTER(u32), C.POINTER(u32)]),
        "dmt_build_light_tree": (u32, [fp, u32, fp]),
        "dmt_build_env_cdf": (C.c_double, [fp, u32, u32, fp, fp]),
        "dmt_build_env_cdf_masked": (C.c_double, [fp, u32, u32, C.POINTER(C.c_uint8), fp, fp]),
        "dmt_env_extract_lights": (u32, [fp, u32, u32, C.c_double, fp, fp, C.POINTER(C.c_uint8)]),
        "dmt_scene_set_moon_map": (st, [vp, fp, u32, u32]),
        "dmt_scene_set_pbr_atlas": (st, [vp, C.POINTER(C.c_uint8), u32, u32]),
        "dmt_scene_get_materials": (st, [vp, fp, C.c_size_t]),
        "dmt_scene_set_materials": (st, [vp, fp, C.c_size_t]),
        "dmt_adjoint": (st, [vp, C.POINTER(FrameParams), vp, vp]),
        "dmt_batch_workload_scene": (st, [C.c_uint64, u32, C.POINTER(GenOptions), C.POINTER(C.c_uint64),
                                          C.POINTER(u32)]),
        "dmt_batch_create": (st, [vp, C.POINTER(vp), C.POINTER(FrameParams), u32, C.POINTER(BatchParams),
                                  C.POINTER(vp)]),
        "dmt_batch_destroy": (None, [vp]),
        "dmt_batch_render": (st, [vp, u32, u32]),
        "dmt_batch_set_params": (st, [vp, C.POINTER(BatchParams)]),
        "dmt_batch_get_params": (st, [vp, C.POINTER(BatchParams)]),
        "dmt_batch_autotune": (st, [vp, u32, C.POINTER(BatchParams), C.POINTER(C.c_float)]),
        "dmt_batch_stats": (st, [vp, C.POINTER(Stats)]),
        "dmt_batch_reset_stats": (st, [vp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def last_error() -> str:
    return lib().dmt_last_error().decode()


def _check(status: int):
    if status != DMT_OK:
        raise DmtError(status, lib().dmt_last_error().decode())


def _fptr(a: np.ndarray | None):
    if a is None or a.size == 0:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _uptr(a: np.ndarray | None):
    if a is None or a.size == 0:
        return None
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


@dataclass
class SceneData:
    """Host copy of a scene in the reference wire formats."""
    prims: np.ndarray = field(default_factory=lambda: np.zeros((0, PRIM_FLOATS), np.float32))
    lights: np.ndarray = field(default_factory=lambda: np.zeros((0, LIGHT_FLOATS), np.float32))
    positions: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.float32))
    indices: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.uint32))
    env: np.ndarray | None = None          # (H, W, 3) float32
    env_intensity: float = 1.0
    frame: FrameParams = field(default_factory=FrameParams)
    env_cdf: tuple | None = None           # (marginal (H,), conditional (H, W), total) given instead of built
    moon: np.ndarray | None = None         # (H, W, 4) moon texture
    sdf: np.ndarray | None = None          # (N, 112) SDF cluster records
    atlas: np.ndarray | None = None        # (256 * tiles, 384, 4) uint8 PBR atlas
    uvs: np.ndarray | None = None          # (V, 2) float32 mesh UVs

    def desc(self) -> SceneDesc:
        for name in ("prims", "lights", "positions"):
            setattr(self, name, np.ascontiguousarray(getattr(self, name), np.float32))
        self.indices = np.ascontiguousarray(self.indices, np.uint32)
        d = SceneDesc()
        d.prims, d.prim_count = _fptr(self.prims), self.prims.shape[0]
        d.lights, d.light_count = _fptr(self.lights), self.lights.shape[0]
        d.positions, d.vertex_count = _fptr(self.positions), self.positions.shape[0]
        d.indices, d.index_count = _uptr(self.indices), self.indices.size
        if self.env is not None:
            self.env = np.ascontiguousarray(self.env, np.float32)
            d.env_rgb, d.env_height, d.env_width = _fptr(self.env), self.env.shape[0], self.env.shape[1]
        d.env_intensity = self.env_intensity
        if self.env_cdf is not None:
            m, c, t = self.env_cdf
            self._cdf = (np.ascontiguousarray(m, np.float32), np.ascontiguousarray(c, np.float32))
            d.env_marginal, d.env_conditional = _fptr(self._cdf[0]), _fptr(self._cdf[1])
            d.env_total_luminance = float(t)
        if self.sdf is not None and np.asarray(self.sdf).size:
            self.sdf = np.ascontiguousarray(self.sdf, np.float32).reshape(-1, SDF_CLUSTER_FLOATS)
            d.sdf_clusters, d.sdf_cluster_count = _fptr(self.sdf), self.sdf.shape[0]
        if self.uvs is not None:
            self.uvs = np.ascontiguousarray(self.uvs, np.float32).reshape(-1, 2)
            d.mesh_uvs = _fptr(self.uvs)
        return d


def generate(config: int, seed: int = 1, subdivisions: int = -1, random_tris: int = -1, spheres: int = -1,
             width: int = 0, height: int = 0, env_width: int = 0, env_height: int = 0, lights: int = -1,
             sdf_clusters: int = -1) -> SceneData:
    """Seeded synthetic workload (dmt_scene_generate), copied to numpy."""
    opt = GenOptions(subdivisions, random_tris, spheres, width, height, env_width, env_height, lights, sdf_clusters)
    out = C.POINTER(GeneratedScene)()
    _check(lib().dmt_scene_generate(config, seed, C.byref(opt), C.byref(out)))
    g = out.contents
    try:
        def arr(ptr, n, dt, shape):
            if n == 0 or not ptr:
                return np.zeros(shape if shape[0] == 0 else (0,) + shape[1:], dt)
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt).reshape(shape).copy()
        sd = SceneData(
            prims=arr(g.prims, g.prim_count * PRIM_FLOATS, np.float32, (g.prim_count, PRIM_FLOATS)),
            lights=arr(g.lights, g.light_count * LIGHT_FLOATS, np.float32, (g.light_count, LIGHT_FLOATS)),
            positions=arr(g.positions, g.vertex_count * 3, np.float32, (g.vertex_count, 3)),
            indices=arr(g.indices, g.index_count, np.uint32, (g.index_count // 3, 3)),
            env=arr(g.env_rgb, g.env_width * g.env_height * 3, np.float32, (g.env_height, g.env_width, 3)),
            env_intensity=g.env_intensity,
            frame=g.frame.copy() if hasattr(g.frame, "copy") else g.frame,
            sdf=arr(g.sdf_clusters, g.sdf_cluster_count * SDF_CLUSTER_FLOATS, np.float32,
                    (g.sdf_cluster_count, SDF_CLUSTER_FLOATS)) if g.sdf_cluster_count else None,
        )
        f = FrameParams()
        C.memmove(C.byref(f), C.byref(g.frame), C.sizeof(FrameParams))
        sd.frame = f
        return sd
    finally:
        lib().dmt_scene_generated_free(out)


def batch_workload_scene(seed: int, index: int) -> tuple[GenOptions, int, int]:
    """Options, scene seed and bounce limit of scene `index` of the batched workload."""
    opt = GenOptions()
    sseed = C.c_uint64()
    mb = C.c_uint32()
    _check(lib().dmt_batch_workload_scene(seed, index, C.byref(opt), C.byref(sseed), C.byref(mb)))
    return opt, sseed.value, mb.value


def generate_batch_scene(seed: int, index: int, scale: float = 1.0) -> SceneData:
    """Scene `index` of the batched workload (config 1 generator); `scale` < 1
    shrinks the resolution and triangle count for quick tests."""
    opt, sseed, mb = batch_workload_scene(seed, index)
    res = max(8, int(opt.width * scale))
    tris = int((20 * 4 ** opt.subdivisions + opt.random_tris) * scale)
    sub = 0
    while sub < 8 and 20 * 4 ** (sub + 1) <= tris:
        sub += 1
    sd = generate(1, sseed, subdivisions=sub, random_tris=max(0, tris - 20 * 4 ** sub), spheres=opt.spheres,
                  width=res, height=res)
    sd.frame.spp = 1
    sd.frame.max_bounces = mb
    return sd


def look_at(eye, target, fov_deg: float, width: int, height: int, base: FrameParams | None = None) -> FrameParams:
    """A frame with the camera at eye looking at target; the other fields
    from base, or zero with the engine's SDF trace defaults (128 steps, 1e-4)."""
    if base is not None:
        f = base.copy()
    else:
        f = FrameParams()
        f.sdf_max_iters, f.sdf_epsilon = 128, 1e-4
    e = (C.c_float * 3)(*eye)
    t = (C.c_float * 3)(*target)
    _check(lib().dmt_camera_look_at(e, t, fov_deg, width, height, C.byref(f)))
    return f


class Device:
    def __init__(self, index: int = 0):
        self._h = C.c_void_p()
        _check(lib().dmt_device_create(index, C.byref(self._h)))

    @property
    def handle(self):
        return self._h

    def stream_ptr(self) -> int:
        s = C.c_void_p()
        _check(lib().dmt_device_stream(self._h, C.byref(s)))
        return s.value or 0

    def synchronize(self):
        _check(lib().dmt_device_synchronize(self._h))

    def is_lost(self) -> bool:
        return bool(lib().dmt_device_is_lost(self._h))

    def close(self):
        if self._h:
            lib().dmt_device_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    def __init__(self, device: Device, data: SceneData, sah_bins: int = 0, max_leaf: int = 0,
                 device_build: bool = False):
        self.device = device
        self.data = data
        self._h = C.c_void_p()
        d = data.desc()
        bp = BuildParams(sah_bins, max_leaf, BUILD_DEVICE if device_build else 0)
        _check(lib().dmt_scene_create(device.handle, C.byref(d), C.byref(bp), C.byref(self._h)))
        if data.moon is not None:
            self.set_moon_map(data.moon)
        if data.atlas is not None:
            self.set_pbr_atlas(data.atlas)

    def set_pbr_atlas(self, rgba8: np.ndarray | None):
        """The PBR atlas: (256 * tiles, 384, 4) uint8, or None."""
        if rgba8 is None:
            _check(lib().dmt_scene_set_pbr_atlas(self._h, None, 0, 0))
            return
        a = np.ascontiguousarray(rgba8, np.uint8)
        _check(lib().dmt_scene_set_pbr_atlas(self._h, a.ctypes.data_as(C.POINTER(C.c_uint8)), a.shape[1], a.shape[0]))

    def set_moon_map(self, rgba: np.ndarray | None):
        if rgba is None:
            _check(lib().dmt_scene_set_moon_map(self._h, None, 0, 0))
            return
        m = np.ascontiguousarray(rgba, np.float32)
        _check(lib().dmt_scene_set_moon_map(self._h, _fptr(m), m.shape[1], m.shape[0]))

    def render(self, frame: FrameParams):
        _check(lib().dmt_render(self._h, C.byref(frame)))

    def render_frames(self, frame: FrameParams, count: int):
        _check(lib().dmt_render_frames(self._h, C.byref(frame), count))

    def readback(self, width: int, height: int) -> np.ndarray:
        out = np.empty((height, width, 4), np.float32)
        _check(lib().dmt_readback_accum(self._h, _fptr(out), out.size))
        return out

    def accum_ptr(self) -> int:
        p = C.c_void_p()
        _check(lib().dmt_accum_device_ptr(self._h, C.byref(p)))
        return p.value or 0

    def update_mesh(self, positions: np.ndarray):
        """New vertex positions of the mesh; the triangle BVH is refit on the device."""
        p = np.ascontiguousarray(positions, np.float32)
        _check(lib().dmt_scene_update_mesh(self._h, _fptr(p), p.size // 3))

    def bind_accum(self, d_rgba: int, width: int, height: int):
        """Render into caller-owned device memory (width*height float4); 0 unbinds."""
        _check(lib().dmt_scene_bind_accum(self._h, C.c_void_p(d_rgba or None), width, height))

    def tonemap(self, d_out: int, exposure: float = 1.0, op: int = 0, mode: int = TONEMAP_DISPLAY):
        """Display image (sRGB of tonemapped mean * exposure, alpha 1), or the
        denoiser colour input (mode LINEAR / HDR), into device memory."""
        p = TonemapParams(exposure, op, mode)
        _check(lib().dmt_tonemap(self._h, C.byref(p), C.c_void_p(d_out)))

    def gbuffer(self, frame: FrameParams, vp_curr, vp_prev, stride: int = 4, **outs):
        """Denoiser inputs into device memory: outs maps depth / motion / normal /
        albedo / specular_albedo / roughness / specular_hit_distance /
        cloud_transmittance to device addresses (stride floats per pixel)."""
        p = GbufferParams()
        p.curr_view_proj[:] = [float(v) for v in np.asarray(vp_curr, np.float32).reshape(16)]
        p.prev_view_proj[:] = [float(v) for v in np.asarray(vp_prev, np.float32).reshape(16)]
        p.pixel_stride = stride
        for k, v in outs.items():
            setattr(p, k, v)
        _check(lib().dmt_render_gbuffer(self._h, C.byref(frame), C.byref(p)))

    def trace(self, origins: np.ndarray, dirs: np.ndarray) -> np.ndarray:
        n = origins.shape[0]
        rays = np.zeros((n, 8), np.float32)
        rays[:, 0:3] = origins
        rays[:, 4:7] = dirs
        out = np.zeros(n, HIT_DTYPE)
        _check(lib().dmt_trace_closest(self._h, _fptr(rays), n, out.ctypes.data_as(C.c_void_p)))
        return out

    def trace_device_noalloc(self, d_rays: int, n: int, d_out: int) -> int:
        """dmt_trace_closest_noalloc on device addresses (n rays of 8 floats
        in, n dmt_hit out); stream-ordered, returns the status code."""
        return lib().dmt_trace_closest_noalloc(self._h, C.c_void_p(d_rays), n, C.c_void_p(d_out))

    def stats(self) -> dict:
        s = Stats()
        _check(lib().dmt_scene_stats(self._h, C.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        _check(lib().dmt_scene_reset_stats(self._h))

    def set_profiling(self, on: bool):
        _check(lib().dmt_scene_set_profiling(self._h, 1 if on else 0))

    def set_schedule(self, wavefront_iterations: int):
        """Per-bounce iterations of a frame before the frame-completion launch
        (SCHEDULE_WAVEFRONT: per-bounce kernels to the end)."""
        _check(lib().dmt_scene_set_schedule(self._h, wavefront_iterations))

    def schedule(self) -> int:
        v = C.c_uint32()
        _check(lib().dmt_scene_get_schedule(self._h, C.byref(v)))
        return v.value

    def measure_traffic(self, frame: FrameParams) -> dict:
        s = Stats()
        _check(lib().dmt_scene_measure_traffic(self._h, C.byref(frame), C.byref(s)))
        return s.as_dict()

    def materials(self) -> np.ndarray:
        n = self.data.prims.shape[0]
        out = np.zeros((n, MATERIAL_FLOATS), np.float32)
        _check(lib().dmt_scene_get_materials(self._h, _fptr(out) if n else None, out.size))
        return out

    def set_materials(self, mats: np.ndarray):
        m = np.ascontiguousarray(mats, np.float32)
        _check(lib().dmt_scene_set_materials(self._h, _fptr(m), m.size))

    def adjoint(self, frame: FrameParams, d_image_grad: int, d_param_grad: int):
        """dLoss/dparams (device prim_count x 7 floats) from dLoss/dI (device W x H float4)."""
        _check(lib().dmt_adjoint(self._h, C.byref(frame), C.c_void_p(d_image_grad), C.c_void_p(d_param_grad)))

    def close(self):
        if self._h:
            lib().dmt_scene_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many scenes rendered by the persistent kernel (dmt_batch_*)."""

    def __init__(self, device: Device, scenes: list, frames: list, params: BatchParams | None = None):
        self.device = device
        self.scenes = list(scenes)   # keeps the Scene objects alive
        n = len(self.scenes)
        hs = (C.c_void_p * n)(*[sc._h.value for sc in self.scenes])
        fs = (FrameParams * n)(*[f.copy() for f in frames])
        self._h = C.c_void_p()
        _check(lib().dmt_batch_create(device.handle, hs, fs, n, C.byref(params) if params else None,
                                      C.byref(self._h)))

    def render(self, first_frame: int, count: int):
        _check(lib().dmt_batch_render(self._h, first_frame, count))

    def set_params(self, p: BatchParams):
        _check(lib().dmt_batch_set_params(self._h, C.byref(p)))

    def params(self) -> BatchParams:
        p = BatchParams()
        _check(lib().dmt_batch_get_params(self._h, C.byref(p)))
        return p

    def autotune(self, frames_per_trial: int = 2) -> tuple[BatchParams, float]:
        p = BatchParams()
        ms = C.c_float()
        _check(lib().dmt_batch_autotune(self._h, frames_per_trial, C.byref(p), C.byref(ms)))
        return p, ms.value

    def stats(self) -> dict:
        s = Stats()
        _check(lib().dmt_batch_stats(self._h, C.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        _check(lib().dmt_batch_reset_stats(self._h))

    def close(self):
        if self._h:
            lib().dmt_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_bvh4(positions: np.ndarray, indices: np.ndarray, bins: int = 0, max_leaf: int = 0):
    """Host SAH BVH4 build (dmt_build_bvh4): returns (nodes float32 (N,32), order uint32, depth)."""
    pos = np.ascontiguousarray(positions, np.float32)
    idx = np.ascontiguousarray(indices, np.uint32).reshape(-1)
    count, depth = C.c_uint32(), C.c_uint32()
    _check(lib().dmt_build_bvh4(_fptr(pos), pos.shape[0], _uptr(idx), idx.size, bins, max_leaf, None, None,
                                C.byref(count), C.byref(depth)))
    nodes = np.zeros((count.value, 32), np.float32)
    order = np.zeros(idx.size // 3, np.uint32)
    _check(lib().dmt_build_bvh4(_fptr(pos), pos.shape[0], _uptr(idx), idx.size, bins, max_leaf,
                                nodes.ctypes.data_as(C.c_void_p), order.ctypes.data_as(C.c_void_p),
                                C.byref(count), C.byref(depth)))
    return nodes, order, depth.value


def build_light_tree(lights: np.ndarray) -> np.ndarray:
    L = np.ascontiguousarray(lights, np.float32)
    n = L.shape[0]
    out = np.zeros((max(2 * n - 1, 1), 16), np.float32)
    cnt = lib().dmt_build_light_tree(_fptr(L), n, _fptr(out))
    return out[:cnt]


def build_env_cdf(env: np.ndarray, mask: np.ndarray | None = None):
    e = np.ascontiguousarray(env, np.float32)
    H, W = e.shape[0], e.shape[1]
    marg = np.zeros(H, np.float32)
    cond = np.zeros(H * W, np.float32)
    if mask is None:
        total = lib().dmt_build_env_cdf(_fptr(e), W, H, _fptr(marg), _fptr(cond))
    else:
        mk = np.ascontiguousarray(mask, np.uint8)
        total = lib().dmt_build_env_cdf_masked(_fptr(e), W, H, mk.ctypes.data_as(C.POINTER(C.c_uint8)), _fptr(marg),
                                               _fptr(cond))
    return marg, cond.reshape(H, W), total


def extract_env_lights(env: np.ndarray, top_frac: float = 0.005):
    """The engine's HDRI light extraction: (count, dir (8, 4), col (8, 4), mask (H, W))."""
    e = np.ascontiguousarray(env, np.float32)
    H, W = e.shape[0], e.shape[1]
    d = np.zeros((8, 4), np.float32)
    c = np.zeros((8, 4), np.float32)
    mask = np.zeros((H, W), np.uint8)
    n = lib().dmt_env_extract_lights(_fptr(e), W, H, top_frac, _fptr(d), _fptr(c),
                                     mask.ctypes.data_as(C.POINTER(C.c_uint8)))
    return int(n), d, c, mask


def set_hdri_lights(frame: FrameParams, count: int, dirs: np.ndarray, cols: np.ndarray):
    frame.hdri_light_count = count
    for i in range(8):
        for a in range(4):
            frame.hdri_dir[i][a] = float(dirs[i][a]) if i < len(dirs) else 0.0
            frame.hdri_col[i][a] = float(cols[i][a]) if i < len(cols) else 0.0
