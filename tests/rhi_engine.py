"""Replays the reference engine's RHI calls for the path tracer against
dmt::rhi::HipDevice (the _dmt_rhi binding).

The engine uploads its scene through the device (prims Engine.cpp:4743-5002
into buffer slot 3, lights :5235-5300 into slot 12, light-tree nodes :5387 into
slot 13, the mesh through CreateBLAS/CreateTLAS :3891-3911, the env map into
storage texture slot 5), then renders each frame with BeginFrame, the binds of
:6867-7217, PushConstants(PtPush) filled at :7821-8676, Dispatch(ceil(w/8),
ceil(h/8), 1) at :9577-9582, Submit and EndFrame.  pack_ptpush writes the
PtPush fields at the byte offsets of the engine's declaration (:7243-7820,
asserted at :9388-9423).
"""
import struct

import numpy as np

import _dmt_rhi as rhi
import dmt

PTPUSH_SIZE = rhi.PTPUSH_SIZE
RGBA8_UNORM, RGBA32F = 1, 4
SAMPLED, STORAGE = 1, 2
NO_TEX_TILE = 0xFFFFFFFF

# PtPush byte offsets (Engine.cpp:7243-7820)
OFF = dict(pos_fovtan=0, fwd_aspect=16, right_xyz=32, up_xyz=48, frame_index=64, reset_accum=68,
           max_bounces=72, tlas_present=76, prim_count=80, spp=84, denoiser_enabled=88, env_map_present=92,
           halton_jitter=96, env_intensity=104, env_total_luminance=108, curr_view_proj=112, prev_view_proj=176,
           sun_and_mode=240, exposure_pad=256,
           w2j_row0=272, w2j_row1=288, w2j_row2=304, dof_params=320, dof_bokeh_params=336, spectral_params=352,
           vol_params=368,
           clouds_p1=384, clouds_p3=416, moon_dir_phase=448, hdri_lights_count=752, write_normal_gbuffer=756,
           write_albedo_gbuffer=760, write_hdr_aux=764, mis_enabled=768, accum_params=784, bvh_params=800,
           sdf_params=816, sdf_params_f=832, write_shadow_vis=856, light_count=880, light_tree_node_count=884,
           light_tree_enabled=888, write_specular_albedo_gbuffer=896, write_roughness_gbuffer=900,
           write_specular_hit_distance_gbuffer=904, restir_enabled=944, smoke_params=960,
           motion_blur_params=1024, mesh_motion_prev=1088, mesh_motion_curr=1104, mesh_tex_indices=1152,
           hosek_params=1184, mat_override_count=1200, water_params0=912, water_params1=928, water_state=1040,
           ocean_params0=1120,
           tonemap_params=1600, sun_extra2=1616, moon_extra=464, sun_extra=480, hdri_lights_dir=496,
           hdri_lights_col=624)


def _offset(name: str) -> int:
    """OFF[name], or hdri_dir<i> / hdri_col<i>: entry i of hdri_lights_dir / _col."""
    if name in OFF:
        return OFF[name]
    for short, field in (("hdri_dir", "hdri_lights_dir"), ("hdri_col", "hdri_lights_col")):
        if name.startswith(short) and name[len(short):].isdigit():
            return OFF[field] + 16 * int(name[len(short):])
    raise KeyError(name)


def pack_ptpush(frame: "dmt.FrameParams", *, prim_count: int, light_count: int, env_present: bool,
                env_intensity: float = 1.0, tlas_present: bool = False, tonemap_op: int = 0, sdf_count: int = 0,
                overrides: dict | None = None) -> bytes:
    """PtPush bytes as the engine fills them for `frame`, with the sun and moon
    below the horizon (the gradient sky without discs) when there is no env map.
    overrides: {field: (struct format, value or tuple)} applied last."""
    b = bytearray(PTPUSH_SIZE)
    put = lambda name, fmt, *v: struct.pack_into("<" + fmt, b, _offset(name), *v)
    put("pos_fovtan", "4f", *frame.pos_fovtan)
    put("fwd_aspect", "4f", *frame.fwd_aspect)
    put("right_xyz", "4f", *frame.right_xyz)
    put("up_xyz", "4f", *frame.up_xyz)
    put("frame_index", "I", frame.frame_index)
    put("reset_accum", "I", frame.reset_accum)
    put("max_bounces", "I", frame.max_bounces)
    put("tlas_present", "I", 1 if tlas_present else 0)
    put("prim_count", "I", prim_count)
    put("spp", "I", frame.spp)
    put("env_map_present", "I", 1 if env_present else 0)
    put("halton_jitter", "2f", 0.25, -0.125)                         # read by the denoiser path only
    put("env_intensity", "f", env_intensity)
    put("sun_and_mode", "4f", 0.0, -0.5, -0.8660254, 1.0 if env_present else 0.0)
    put("w2j_row0", "4f", 1.0, 0.0, 0.0, 1.0)                         # .w = r_hdr_pipeline
    put("w2j_row1", "4f", 0.0, 1.0, 0.0, float(frame.refract_bounces))
    put("w2j_row2", "4f", 0.0, 0.0, 1.0, frame.firefly_clamp)
    put("moon_dir_phase", "4f", 0.3, -0.6, 0.74, 1.5)
    put("mis_enabled", "I", frame.mis_enabled)
    put("accum_params", "4f", frame.ema_alpha, 0.0, 0.0, 0.0)
    put("light_count", "I", light_count)
    put("light_tree_node_count", "I", max(0, 2 * light_count - 1) if frame.light_tree_enabled else 0)
    put("light_tree_enabled", "I", 1 if (frame.light_tree_enabled and light_count) else 0)
    put("mesh_tex_indices", "4I", *([NO_TEX_TILE] * 4))
    put("dof_params", "4f", *frame.dof)
    put("dof_bokeh_params", "4f", frame.cat_eye, 0.0, 0.0, 0.0)
    put("spectral_params", "4f", 1.0 if frame.spectral else 0.0, frame.cauchy_c, 0.0, 0.0)
    put("motion_blur_params", "4f", 1.0 if frame.motion_blur else 0.0, 1.0 / 60.0, 0.0, 0.0)
    put("mesh_motion_prev", "4f", *frame.mesh_shift_prev, 1.0 if frame.mesh_shift_active else 0.0)
    put("mesh_motion_curr", "4f", *frame.mesh_shift_curr, 0.0)
    put("hosek_params", "4f", 2.5, 0.3, 1.2, 0.0)                      # read by sky mode 3 only
    # SDF clusters at buffer slot 10, the frame's step cap and surface distance,
    # the fractal knobs at the engine's defaults (Engine.cpp:7947-7981)
    put("sdf_params", "4I", sdf_count, frame.sdf_max_iters, 0, 12)
    put("sdf_params_f", "4f", frame.sdf_epsilon, 8.0, 1.0, 0.0)
    put("tonemap_params", "4I", tonemap_op, 0, 0, 0)
    put("sun_extra2", "4f", 1.0, 0.0, 0.0, 0.0)
    for name, (fmt, v) in (overrides or {}).items():
        put(name, fmt, *(v if isinstance(v, tuple) else (v,)))
    return bytes(b)


def rgba_env(env: np.ndarray) -> np.ndarray:
    """(H, W, 3) linear RGB -> (H, W, 4) RGBA32F, as the engine uploads it."""
    h, w, _ = env.shape
    out = np.ones((h, w, 4), np.float32)
    out[..., :3] = env
    return out


class EngineSim:
    """One scene uploaded the engine's way, rendered frame after frame."""

    def __init__(self, dev: "rhi.HipDevice", sd: "dmt.SceneData", exposure: float = 1.0):
        self.dev, self.sd = dev, sd
        f = sd.frame
        self.width, self.height = f.width, f.height
        self.prims = dev.create_buffer(max(112, sd.prims.nbytes))
        dev.write_buffer(self.prims, np.ascontiguousarray(sd.prims, np.float32))
        self.lights = dev.create_buffer(max(80, sd.lights.nbytes))
        dev.write_buffer(self.lights, np.ascontiguousarray(sd.lights, np.float32))
        # light-tree nodes (64 B each) and env CDFs: uploaded by the engine, not read by this backend
        self.tree = dev.create_buffer(64 * max(1, 2 * sd.lights.shape[0]))
        self.exposure = dev.create_buffer(16)
        dev.write_buffer(self.exposure, np.array([exposure, 0, 0, 0], np.float32))
        self.tlas = 0
        if sd.indices.size:
            blas = dev.create_blas(sd.positions, sd.indices)
            assert blas, dev.last_error()
            self.tlas = dev.create_tlas([(blas, None, 0, 0xFF)])
        self.env = 0
        self.env_present = sd.env is not None and sd.env.size > 0
        if self.env_present:
            h, w, _ = sd.env.shape
            self.env = dev.create_texture(w, h, RGBA32F, SAMPLED)
            assert dev.write_texture(self.env, rgba_env(sd.env))
        self.accum = dev.create_texture(self.width, self.height, RGBA32F, STORAGE)
        assert self.accum, dev.last_error()
        self.sdf = 0
        self.sdf_count = 0 if sd.sdf is None else int(np.asarray(sd.sdf).reshape(-1, 112).shape[0])
        if self.sdf_count:
            self.sdf = dev.create_buffer(self.sdf_count * 448)
            dev.write_buffer(self.sdf, np.ascontiguousarray(sd.sdf, np.float32))
        self.pipeline = dev.create_pipeline("pathtrace")
        dev.resize(self.width, self.height)

    def push(self, frame: "dmt.FrameParams", **kw) -> bytes:
        return pack_ptpush(frame, prim_count=self.sd.prims.shape[0], light_count=self.sd.lights.shape[0],
                           env_present=self.env_present, env_intensity=self.sd.env_intensity,
                           tlas_present=bool(self.tlas), sdf_count=self.sdf_count, **kw)

    def frame(self, frame: "dmt.FrameParams", push: bytes | None = None, textures: dict | None = None,
              buffers: dict | None = None, **kw):
        """One engine frame; `textures` / `buffers` bind further storage
        textures and buffers {slot: handle}."""
        d = self.dev
        out, w, h, _ = d.begin_frame()
        d.bind_pipeline(self.pipeline)
        d.bind_texture(0, out)
        d.bind_texture(1, self.accum)
        if self.tlas:
            d.bind_accel(2, self.tlas)
        d.bind_buffer(3, self.prims)
        if self.env:
            d.bind_texture(5, self.env)
        d.bind_buffer(6, self.exposure)
        if self.sdf:
            d.bind_buffer(10, self.sdf)
        d.bind_buffer(12, self.lights)
        d.bind_buffer(13, self.tree)
        for slot, tex in (textures or {}).items():
            d.bind_texture(slot, tex)
        for slot, buf in (buffers or {}).items():
            d.bind_buffer(slot, buf)
        d.push_constants(push if push is not None else self.push(frame, **kw))
        d.dispatch((w + 7) // 8, (h + 7) // 8, 1)
        d.submit()
        d.end_frame()
        return out

    def frames(self, first: "dmt.FrameParams", count: int, **kw):
        out = None
        for k in range(count):
            f = first.copy()
            f.frame_index = first.frame_index + k
            f.reset_accum = 1 if (k == 0 and first.reset_accum) else 0
            out = self.frame(f, **kw)
        return out

    def accum_image(self) -> np.ndarray:
        return self.dev.readback_texture(self.accum)


def tonemap_reference(mean_rgb: np.ndarray, exposure: float, op: int, oetf: bool = True) -> np.ndarray:
    """The display transform in float64 (SoftwareTracer.cpp:263-341, 995-1007):
    sRGB OETF of tonemap(mean * exposure); without the OETF when oetf is false."""
    c = mean_rgb.astype(np.float64) * exposure
    if op == 1:
        ir = np.array([[0.842479062253094, 0.0784335999999992, 0.0792237451477643],
                       [0.0423282422610123, 0.878468636469772, 0.0791661274605434],
                       [0.0423756549057051, 0.0784336, 0.879142973793104]])
        v = c @ ir.T
        lo, hi = -12.47393, 4.026069
        v = (np.clip(np.log2(np.maximum(v, 1e-10)), lo, hi) - lo) / (hi - lo)
        v = (15.5 * v ** 6 - 40.14 * v ** 5 + 31.96 * v ** 4 - 6.868 * v ** 3 + 0.4298 * v ** 2
             + 0.1191 * v - 0.00232)
        orr = np.array([[1.19687900512017, -0.0980208811401368, -0.0990297440797205],
                        [-0.0528968517574562, 1.15190312990417, -0.0989611768448433],
                        [-0.0529716355144438, -0.0980434501171241, 1.15107367264116]])
        v = v @ orr.T
        t = np.clip(np.maximum(v, 0.0) ** 2.2, 0.0, 1.0)
    elif op == 2:
        np.seterr(divide="ignore", invalid="ignore")   # the unused branch of np.where at peak 0
        start, desat = 0.8 - 0.04, 0.15
        x = c.min(axis=-1, keepdims=True)
        offset = np.where(x < 0.08, x - 6.25 * x * x, 0.04)
        c = c - offset
        peak = c.max(axis=-1, keepdims=True)
        d = 1.0 - start
        new_peak = 1.0 - d * d / (peak + d - start)
        g = 1.0 - 1.0 / (desat * (peak - new_peak) + 1.0)
        comp = c * (new_peak / peak)
        comp = comp * (1.0 - g) + new_peak * g
        t = np.where(peak < start, c, comp)
    elif op == 3:
        t = np.clip(c / (1.0 + c), 0.0, 1.0)
    elif op == 4:
        t = np.clip(c, 0.0, 1.0)
    else:
        t = np.clip((c * (2.51 * c + 0.03)) / (c * (2.43 * c + 0.59) + 0.14), 0.0, 1.0)
    if not oetf:
        return t
    return np.where(t <= 0.0031308, t * 12.92, 1.055 * np.power(np.maximum(t, 0.0), 1.0 / 2.4) - 0.055)


def view_proj(frame, pos=None, near=0.05, far=500.0):
    """Column-major (glm::value_ptr order) proj * view of the frame's camera in
    float64: the right-handed look-at of the camera basis and the zero-to-one
    depth perspective with the frame's tan(fov/2) and aspect (the engine's
    glm::lookAtRH / perspectiveRH_ZO, Engine.cpp:7239-7241)."""
    p = np.array(frame.pos_fovtan[:3], np.float64) if pos is None else np.asarray(pos, np.float64)
    f = np.array(frame.fwd_aspect[:3], np.float64)
    r = np.array(frame.right_xyz[:3], np.float64)
    u = np.array(frame.up_xyz[:3], np.float64)
    V = np.eye(4)
    V[0, :3], V[1, :3], V[2, :3] = r, u, -f
    V[:3, 3] = -V[:3, :3] @ p
    t, a = float(frame.pos_fovtan[3]), float(frame.fwd_aspect[3])
    P = np.zeros((4, 4))
    P[0, 0], P[1, 1] = 1.0 / (a * t), 1.0 / t
    P[2, 2], P[2, 3], P[3, 2] = far / (near - far), near * far / (near - far), -1.0
    return (P @ V).T.reshape(16)


def pixel_dirs(frame):
    """Pixel-centre camera directions in float64."""
    w, h = frame.width, frame.height
    x, y = np.meshgrid(np.arange(w) + 0.5, np.arange(h) + 0.5)
    ux = x / w * 2.0 - 1.0
    uy = -(y / h * 2.0 - 1.0)
    t, a = frame.pos_fovtan[3], frame.fwd_aspect[3]
    d = (np.array(frame.fwd_aspect[:3]) + np.multiply.outer(ux * a * t, frame.right_xyz[:3])
         + np.multiply.outer(uy * t, frame.up_xyz[:3]))
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def to_pixel(M, pts, w, h):
    M = M.reshape(4, 4).T
    c = pts @ M[:, :3].T + M[:, 3]
    cw = np.maximum(c[..., 3], 1e-6)
    return np.stack([(c[..., 0] / cw * 0.5 + 0.5) * w, (-c[..., 1] / cw * 0.5 + 0.5) * h], axis=-1)
