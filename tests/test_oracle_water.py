"""The oracle's water (PathTrace.slang:5089-5296 transmittance and
waterShadingNormal, :5847-5872 the camera's medium, :6014-6052 Beer's law and
the deep-water tint, :6768-6906 the water surface; PathTraceCloud.slang:31-69
cloud_value_noise), pinned by closed forms: the value noise equals the
lowbias32 hash at lattice points and is continuous across cells, the wave
normal is the macro normal bent by the central-difference gradient of its two
noise octaves, a camera in the water sees a lit floor dimmed by exp(-a d) on
both the camera and the shadow segment and escaping rays return the deep
tint, a flat surface reflects with Schlick's F0 0.02 and refracts by Snell's
law into absorbing water, and a shadow ray leaving the water carries
exp(-a s) (1 - F)."""
import math

import numpy as np
import pytest

import dmt
import oracle_ctypes as orc
from scenes import black_env, frame, plane, point_light, scene

MAT_WATER = dmt.MAT_WATER


def _lowbias32(x):
    x = np.uint32(x)
    x ^= x >> np.uint32(16)
    x = np.uint32((int(x) * 0x7FEB352D) & 0xFFFFFFFF)
    x ^= x >> np.uint32(15)
    x = np.uint32((int(x) * 0x846CA68B) & 0xFFFFFFFF)
    x ^= x >> np.uint32(16)
    return int(x)


def _lattice_hash(i, j, k):
    h = _lowbias32(i & 0xFFFFFFFF)
    h = _lowbias32(h ^ (j & 0xFFFFFFFF))
    h = _lowbias32(h ^ (k & 0xFFFFFFFF))
    return (h & 0xFFFFFF) / 16777216.0


def test_value_noise_is_the_hash_at_lattice_points_and_continuous():
    # the published lowbias32 constants: 0 -> 0, and two known values
    assert _lowbias32(0) == 0
    rng = np.random.default_rng(4)
    pts = rng.integers(-300, 300, size=(64, 3))
    got = orc.value_noise(pts.astype(np.float32))
    want = np.array([_lattice_hash(int(a), int(b), int(c)) for a, b, c in pts], np.float32)
    assert np.array_equal(got, want)
    # continuous across the cell faces, and in [0, 1)
    base = rng.uniform(-50, 50, size=(256, 3)).astype(np.float32)
    face = np.floor(base) + np.float32(1.0)
    lo = face - np.float32(1e-4)
    hi = face + np.float32(1e-4)
    for axis in range(3):
        a, b = base.copy(), base.copy()
        a[:, axis], b[:, axis] = lo[:, axis], hi[:, axis]
        assert np.max(np.abs(orc.value_noise(a) - orc.value_noise(b))) <= 2e-4
    v = orc.value_noise(rng.uniform(-1e3, 1e3, size=(4096, 3)))
    assert v.min() >= 0.0 and v.max() < 1.0 and 0.4 < v.mean() < 0.6


def _wave_frame(scale=1.3, amp=0.4, speed=0.8, t=2.5):
    f = dmt.FrameParams()
    f.wave[:] = (scale, amp, speed, t)
    return f


def test_wave_normal_is_the_bent_macro_normal():
    rng = np.random.default_rng(5)
    p = rng.uniform(-20, 20, size=(128, 3)).astype(np.float32)
    up = np.tile(np.float32([0, 1, 0]), (128, 1))
    # no amplitude or no scale: the macro normal itself
    for f in (_wave_frame(amp=0.0), _wave_frame(scale=0.0), _wave_frame(amp=-1.0)):
        assert np.array_equal(orc.water_normal(f, p, up), up)
    f = _wave_frame()
    scale, amp, speed, t = f.wave
    n = orc.water_normal(f, p, up).astype(np.float64)
    assert np.max(np.abs(np.linalg.norm(n, axis=1) - 1)) <= 2e-6
    # the height field: octave A drifting (0.13, 0.07) t', octave B at 2.02x
    # drifting (-0.09, 0.11) t', half weight; central differences, step 0.15
    tt = np.float32(t) * np.float32(speed)
    eps = np.float32(0.15)

    def hgt(x, z, x2, z2):
        a = orc.value_noise(np.stack([x + np.float32(0.13) * tt, np.zeros_like(x), z + np.float32(0.07) * tt], 1))
        b = orc.value_noise(np.stack([x2 + np.float32(-0.09) * tt, np.zeros_like(x), z2 + np.float32(0.11) * tt], 1))
        return a.astype(np.float64) + 0.5 * b

    px, pz = p[:, 0] * np.float32(scale), p[:, 2] * np.float32(scale)
    k = np.float32(2.02)
    dhx = (hgt(px + eps, pz, px * k + eps, pz * k) - hgt(px - eps, pz, px * k - eps, pz * k)) / 0.3
    dhz = (hgt(px, pz + eps, px * k, pz * k + eps) - hgt(px, pz - eps, px * k, pz * k - eps)) / 0.3
    # for n = +y the basis is tx = normalize(cross(x, y)) = +z... the frame
    # picks t1 = x (|n.y| >= 0.99): tx = cross(x, y) = z, tz = cross(y, z) = x
    want = np.stack([-amp * dhz, np.ones_like(dhx), -amp * dhx], 1)
    want /= np.linalg.norm(want, axis=1, keepdims=True)
    assert np.max(np.abs(n - want)) <= 2e-5
    # a tilted macro normal: the bend stays orthogonal to it
    m = np.tile(np.float32([0.6, 0.8, 0.0]), (128, 1))
    nt = orc.water_normal(f, p, m).astype(np.float64)
    bend = nt / np.sum(nt * m, axis=1, keepdims=True) - m
    assert np.max(np.abs(np.sum(bend * m, axis=1))) <= 2e-5
    assert np.max(np.linalg.norm(bend, axis=1)) > 1e-2


def _fixed_centre(f):
    f.flags |= dmt.DMT_FRAME_FIXED_JITTER
    f.jitter[:] = (0.0, 0.0)
    return f


def _centre_dirs(f):
    w, h = f.width, f.height
    fwd = np.array(f.fwd_aspect[:3]); rt = np.array(f.right_xyz[:3]); up = np.array(f.up_xyz[:3])
    tanf, aspect = f.pos_fovtan[3], f.fwd_aspect[3]
    ys, xs = np.mgrid[0:h, 0:w]
    ux = ((xs + 0.5) / w) * 2 - 1
    uy = -(((ys + 0.5) / h) * 2 - 1)
    d = fwd + rt * (ux * aspect * tanf)[..., None] + up * (uy * tanf)[..., None]
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def test_camera_in_water_sees_a_floor_dimmed_on_both_segments():
    albedo, I = 0.6, 20.0
    Lp = np.array([0.5, 1.0, -0.3])
    sd = scene([plane((0, 1, 0), 2.0, albedo=(albedo,) * 3)], [point_light(Lp, (I, I, I))], env=black_env())
    f = _fixed_centre(frame((0, 0.5, 4), (0, -2, 0), 30, 8, 6, spp=1, bounces=2))
    a = np.array([0.05, 0.1, 0.2])
    f.camera_in_water = 1
    f.water_absorb[:] = a
    img, cnt = orc.OracleScene.from_scene_data(sd).render(f, 32, 0)
    o = np.array(f.pos_fovtan[:3])
    d = _centre_dirs(f)
    t = -(o[1] + 2.0) / d[..., 1]
    P = o + d * t[..., None]
    v = Lp - P
    dist = np.linalg.norm(v, axis=-1)
    cos = v[..., 1] / dist
    want = (np.exp(-a * t[..., None]) * (I / dist ** 2)[..., None] * np.exp(-a * (dist - 1e-3)[..., None]) *
            (albedo / math.pi) * cos[..., None])
    assert np.max(np.abs(img[..., :3] / want - 1)) <= 3e-5
    # nothing in view: the deep-water tint, exactly (camera segment, unclamped)
    sd2 = scene([], [], env=black_env())
    f.deep_color[:] = (0.1, 0.2, 0.3)
    img2, _ = orc.OracleScene.from_scene_data(sd2).render(f, 32, 0)
    assert np.all(img2[..., :3] == np.float32([0.1, 0.2, 0.3]))


@pytest.mark.parametrize("theta_deg", [20.0, 60.0, 75.0])
def test_flat_surface_reflects_with_schlick_and_refracts_by_snell(theta_deg):
    n_w, a_g, depth = 1.33, 0.3, 3.0
    water = plane((0, 1, 0), 0.0, albedo=(1, 1, 1), mat=MAT_WATER, ior=n_w)
    sky = plane((0, 1, 0), -5.0, albedo=(0, 0, 0), emission=(1, 0, 0))
    floor = plane((0, 1, 0), depth, albedo=(0, 0, 0), emission=(0, 1, 0))
    sd = scene([water, sky, floor], [], env=black_env())
    th = math.radians(theta_deg)
    f = frame((0, 2, 0), (2 * math.tan(th), 0, 0), 0.01, 64, 64, spp=32, bounces=2)
    f.water_absorb[:] = (0.0, a_g, 0.0)
    img, _ = orc.OracleScene.from_scene_data(sd).render(f, 32, 0)
    cos_i = math.cos(th)
    F = 0.02 + 0.98 * (1 - cos_i) ** 5
    sin_t = math.sin(th) / n_w
    cos_t = math.sqrt(1 - sin_t * sin_t)
    red, green = img[..., 0].mean(), img[..., 1].mean()
    N = 64 * 64 * 32
    sig = math.sqrt(F * (1 - F) / N)
    assert abs(red - F) <= 5 * sig + 1e-6
    assert abs(green - (1 - F) * math.exp(-a_g * depth / cos_t)) <= 5 * sig * math.exp(-a_g * depth / cos_t) + 1e-6


def test_shadow_ray_leaving_the_water_carries_beer_and_fresnel():
    albedo, I, n_w = 0.6, 30.0, 1.33
    a = np.array([0.07, 0.12, 0.25])
    Lp = np.array([0.7, 3.0, -0.4])
    water = plane((0, 1, 0), 0.0, albedo=(1, 1, 1), mat=MAT_WATER, ior=n_w)
    floor = plane((0, 1, 0), 2.0, albedo=(albedo,) * 3)
    sd = scene([water, floor], [point_light(Lp, (I, I, I))], env=black_env())
    f = _fixed_centre(frame((0, -0.5, 3), (0, -2, 0), 30, 8, 6, spp=1, bounces=1))
    f.camera_in_water = 1
    f.water_absorb[:] = a
    img, _ = orc.OracleScene.from_scene_data(sd).render(f, 32, 0)
    o = np.array(f.pos_fovtan[:3])
    d = _centre_dirs(f)
    t = -(o[1] + 2.0) / d[..., 1]
    P = o + d * t[..., None]
    v = Lp - P
    dist = np.linalg.norm(v, axis=-1)
    u = v / dist[..., None]
    s = (0.0 - (P[..., 1] + 1e-3)) / u[..., 1]          # to the surface, from the offset origin
    cos_i = u[..., 1]
    F = 0.02 + 0.98 * (1 - cos_i) ** 5
    sin_t = n_w * np.sqrt(1 - cos_i ** 2)
    assert np.all(sin_t < 1)                             # no total internal reflection in view
    trans = np.exp(-a * s[..., None]) * (1 - F)[..., None]
    want = np.exp(-a * t[..., None]) * (I / dist ** 2)[..., None] * trans * (albedo / math.pi) * u[..., 1:2]
    assert np.max(np.abs(img[..., :3] / want - 1)) <= 3e-5
    # the same floor without the surface above: no Fresnel factor, Beer's law
    # along the whole shadow segment
    sd2 = scene([floor], [point_light(Lp, (I, I, I))], env=black_env())
    img2, _ = orc.OracleScene.from_scene_data(sd2).render(f, 32, 0)
    want2 = (np.exp(-a * t[..., None]) * (I / dist ** 2)[..., None] * np.exp(-a * (dist - 1e-3)[..., None]) *
             (albedo / math.pi) * u[..., 1:2])
    assert np.max(np.abs(img2[..., :3] / want2 - 1)) <= 3e-5
