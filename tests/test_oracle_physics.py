"""Closed-form checks of the oracle's light transport (CPU only).

Each test sets up a scene whose expected radiance follows from the
mathematics, and checks the oracle against it:
  * point light: radiance = albedo/pi * I cos(theta) / d^2, deterministic;
  * spot light: the same times the linear cone ramp;
  * quad light: albedo/pi * L * E_polygon (Lambert's polygon formula);
  * sphere light: albedo/pi * L * pi (R/d)^2 cos(theta) for a fully visible sphere;
  * white furnace: a Lambert sphere of albedo a under a uniform env L=1
    reflects exactly a (env NEE + BSDF sampling under MIS, PathTrace.slang:6457-6492,
    6056-6085; the furnace rationale is stated at PathTrace.slang:6536-6541);
  * glass furnace: a clear dielectric leaves throughput unchanged
    (PathTrace.slang:6709-6767), so every escaping path carries L=1 exactly;
  * mirror: a smooth metal plane returns albedo * Le of an emitter behind the
    camera, exactly;
  * closed emitter: inside an emissive Lambert sphere every vertex adds T*Le,
    so L is the bounce series Le (1 - a^N) / (1 - a) without Russian roulette
    and its expectation with it;
  * accumulation: the running mean of K frames and the EMA recursion.
Wrong pdfs (area->solid angle), a dropped 1/pi, a wrong MIS weight, a wrong
Fresnel branch, a wrong reflection, a dropped 1/p of Russian roulette or a
wrong accumulation weight each fail one of them.
"""
import math

import numpy as np
import pytest

import oracle_ctypes as orc
import scenes


def _probe(sd, eye, target, spp=1, frames=1, w=4, h=4, fov=0.02, precision=64, bounces=2):
    f = scenes.frame(eye, target, fov, w, h, spp=spp, bounces=bounces)
    sc = orc.OracleScene.from_scene_data(sd)
    acc = None
    for k in range(frames):
        f.frame_index = k
        f.reset_accum = 1 if k == 0 else 0
        acc, _ = sc.render(f, precision, 0, accum=acc)
    return acc[..., :3].reshape(-1, 3).mean(axis=0)


def _floor(albedo=(0.6, 0.5, 0.4)):
    return scenes.plane((0, 1, 0), 0.0, albedo=albedo)


def test_point_light_is_exact():
    a = np.array([0.6, 0.5, 0.4])
    I = np.array([3.0, 2.0, 1.0])
    lp = np.array([0.7, 2.0, -0.3])
    sd = scenes.scene([_floor(a)], [scenes.point_light(lp, I)], env=scenes.black_env())
    x = np.array([0.0, 0.0, 0.0])
    got = _probe(sd, (0.0, 1.0, 1.5), x, precision=64)
    d = lp - x
    expect = a / math.pi * I * (d[1] / np.linalg.norm(d)) / float(d @ d)
    assert np.allclose(got, expect, rtol=2e-4)


def test_spot_light_ramp():
    a = np.array([0.5, 0.5, 0.5])
    I = np.array([4.0, 4.0, 4.0])
    lp = np.array([0.0, 2.0, 0.0])
    # axis tilted so the probe point sits inside the linear ramp
    ax = np.array([0.3, -1.0, 0.0]); ax /= np.linalg.norm(ax)
    L = scenes.light(1, lp, I, dir=ax, cos_outer=0.90, cos_inner=0.99)
    sd = scenes.scene([_floor(a)], [L], env=scenes.black_env())
    x = np.array([0.0, 0.0, 0.0])
    got = _probe(sd, (0.0, 1.0, 1.5), x)
    d = lp - x
    dist = np.linalg.norm(d)
    cd = float(np.dot(-d / dist, ax))
    ramp = 1.0 if cd >= 0.99 else max(0.0, (cd - 0.90) / (0.99 - 0.90))
    expect = a / math.pi * I * ramp * (d[1] / dist) / dist ** 2
    assert 0.0 < ramp < 1.0
    assert np.allclose(got, expect, rtol=2e-4)


def test_quad_light_irradiance():
    a = np.array([0.8, 0.8, 0.8])
    Le = 5.0
    c = np.array([0.0, 2.0, 0.0])
    u = np.array([0.6, 0.0, 0.0])
    vh = 0.4
    n = np.array([0.0, -1.0, 0.0])
    sd = scenes.scene([_floor(a)], [scenes.quad_light(c, n, u, vh, (Le, Le, Le))], env=scenes.black_env())
    for x in (np.array([0.0, 0.0, 0.0]), np.array([1.2, 0.0, 0.5])):
        got = _probe(sd, x + np.array([0.0, 1.0, 1.5]), x, spp=32, frames=40)
        v_axis = np.cross(n, u); v_axis /= np.linalg.norm(v_axis)
        verts = [c - u - v_axis * vh, c + u - v_axis * vh, c + u + v_axis * vh, c - u + v_axis * vh]
        E = scenes.polygon_irradiance(x, np.array([0.0, 1.0, 0.0]), verts)
        expect = a / math.pi * Le * E
        assert np.allclose(got, expect, rtol=0.02), (x, got, expect)


def test_triangle_light_irradiance():
    """Emissive triangle (light type 4) over a Lambert floor: the radiance
    equals albedo / pi * Le * Lambert's polygon irradiance of the triangle."""
    a = np.array([0.8, 0.7, 0.6])
    Le = 4.0
    A, B, C = np.array([-0.5, 2.0, -0.3]), np.array([0.7, 2.2, 0.1]), np.array([0.1, 1.8, 0.8])
    if np.cross(B - A, C - A)[1] > 0:   # face the floor
        B, C = C, B
    sd = scenes.scene([_floor(a)], [scenes.triangle_light(A, B, C, (Le, Le, Le))], env=scenes.black_env())
    for x in (np.array([0.0, 0.0, 0.0]), np.array([1.0, 0.0, -0.7])):
        got = _probe(sd, x + np.array([0.0, 1.0, 1.5]), x, spp=32, frames=40)
        E = scenes.polygon_irradiance(x, np.array([0.0, 1.0, 0.0]), [A, B, C])
        expect = a / math.pi * Le * E
        assert np.allclose(got, expect, rtol=0.02), (x, got, expect)
    # one-sided: flipped, the triangle lights nothing below it
    sd = scenes.scene([_floor(a)], [scenes.triangle_light(A, C, B, (Le, Le, Le))], env=scenes.black_env())
    got = _probe(sd, (0.0, 1.0, 1.5), (0.0, 0.0, 0.0), spp=8, frames=4)
    assert np.all(got == 0.0)


def test_sphere_light_irradiance():
    a = np.array([0.7, 0.7, 0.7])
    Le, R = 3.0, 0.3
    lp = np.array([0.4, 2.5, 0.0])
    sd = scenes.scene([_floor(a)], [scenes.sphere_light(lp, R, (Le, Le, Le))], env=scenes.black_env())
    x = np.array([0.0, 0.0, 0.0])
    got = _probe(sd, (0.0, 1.0, 1.5), x, spp=32, frames=40)
    d = lp - x
    dist = np.linalg.norm(d)
    E = Le * math.pi * (R / dist) ** 2 * (d[1] / dist)
    assert np.allclose(got, a / math.pi * E, rtol=0.02)


def test_white_furnace_lambert_sphere():
    a = 0.5
    env = np.ones((32, 64, 3), np.float32)
    sd = scenes.scene([scenes.sphere((0, 0, 0), 1.0, albedo=(a, a, a))], [], env=env)
    got = _probe(sd, (0.0, 0.0, 4.0), (0.0, 0.0, 0.0), spp=16, frames=20, w=16, h=16, fov=20.0, bounces=4)
    assert np.allclose(got, a, rtol=0.01), got


def test_glass_furnace_preserves_radiance():
    env = np.ones((16, 32, 3), np.float32)
    sd = scenes.scene([scenes.sphere((0, 0, 0), 1.0, albedo=(1, 1, 1), mat=2, ior=1.5)], [], env=env)
    f = scenes.frame((0, 0, 4), (0, 0, 0), 25.0, 32, 32, spp=4, bounces=32)
    img, _ = orc.OracleScene.from_scene_data(sd).render(f, 32, 0)
    rgb = img[..., :3]
    # paths through clear glass keep T = 1 and end in L = 1; only paths cut
    # by the 32-bounce limit (internal reflections) come back darker
    assert np.mean(rgb == 1.0) > 0.97
    assert np.all(rgb <= 1.0)


def test_mirror_reflects_an_emitter_exactly():
    """A smooth metal plane facing the camera sends the centre rays straight
    back to an emissive sphere behind the camera: radiance = albedo * Le
    (PathTrace.slang:6683-6708: T *= albedo, r = reflect(rd, n) at roughness
    0; the emitter adds T * Le at bounce 1, :6193-6199)."""
    albedo, Le = np.float32(0.5), np.float32(1.5)
    mirror = scenes.plane((0, 0, 1), 0.0, albedo=(albedo, albedo, albedo), mat=1, rough=0.0)
    emitter = scenes.sphere((0, 0, 9), 1.0, albedo=(0, 0, 0), emission=(Le, Le, Le))
    sd = scenes.scene([mirror, emitter], [], env=np.zeros((8, 16, 3), np.float32))
    f = scenes.frame((0.0, 0.0, 5.0), (0.0, 0.0, 0.0), 10.0, 16, 16, bounces=3)
    img, _ = orc.OracleScene.from_scene_data(sd).render(f, 32, 0)
    rgb = img[..., :3]
    hit = np.all(rgb == albedo * Le, axis=-1)
    assert np.all(hit | np.all(rgb == 0.0, axis=-1))   # each pixel sees the emitter or nothing
    assert hit[6:10, 6:10].all() and 40 < hit.sum() < 256


def _closed_emitter(albedo, Le, rr, bounces, frames, spp=4):
    """Camera inside an emissive Lambert sphere, black env, no lights: every
    vertex adds T * Le and multiplies T by the albedo."""
    inside = scenes.sphere((0, 0, 0), 20.0, albedo=(albedo,) * 3, emission=(Le,) * 3)
    sd = scenes.scene([inside], [], env=np.zeros((8, 16, 3), np.float32))
    f = scenes.frame((0.0, 0.0, 0.0), (0.0, 0.0, -1.0), 60.0, 24, 16, spp=spp, bounces=bounces)
    f.flags = 0 if rr else 1
    sc = orc.OracleScene.from_scene_data(sd)
    acc = None
    for k in range(frames):
        f.frame_index = k
        f.reset_accum = 1 if k == 0 else 0
        acc, _ = sc.render(f, 64, 0, accum=acc)
    return acc[..., :3]


def test_closed_emitter_sums_the_bounce_series():
    """Without Russian roulette each path hits the enclosure at every vertex:
    L = Le (1 + a + ... + a^(N-1)) on every pixel (the max_bounces limit and
    the throughput chain, PathTrace.slang:6193-6199, 6668)."""
    a, Le, n = 0.7, 0.5, 8
    img = _closed_emitter(a, Le, rr=False, bounces=n, frames=1)
    assert np.allclose(img, Le * (1.0 - a ** n) / (1.0 - a), rtol=1e-6)   # fp64 render, fp32 accumulator


def test_russian_roulette_is_unbiased():
    """From bounce 2 a path survives with p = clamp(max(T), 0.05, 1) and
    carries T / p (PathTrace.slang:6908-6913): the mean stays the bounce
    series (without the 1/p it would fall short of it)."""
    a, Le, n = 0.7, 0.5, 8
    want = Le * (1.0 - a ** n) / (1.0 - a)
    img = _closed_emitter(a, Le, rr=True, bounces=n, frames=24)
    assert abs(img.mean() - want) < 0.01 * want
    assert img.std() > 0.01   # the estimator is random, pixel by pixel


def test_accumulation_is_the_running_mean_and_the_ema():
    """accum after K frames = the mean of the K frame values; with
    r_accum_ema_alpha = e it follows accum = e accum + (1 - e) v from the
    second frame on (PathTrace.slang:7736-7781)."""
    sd = scenes.scene([scenes.sphere((0, 0, 0), 1.0, albedo=(0.6, 0.5, 0.4)), _floor()],
                      [scenes.point_light((1, 4, 2), (30, 30, 30))])
    f = scenes.frame((0.0, 1.0, 5.0), (0.0, 0.0, 0.0), 40.0, 12, 8, spp=2, bounces=4)
    sc = orc.OracleScene.from_scene_data(sd)
    singles = []
    for k in range(5):
        g = f.copy()
        g.frame_index, g.reset_accum = k, 1
        img, _ = sc.render(g, 32, 0)
        singles.append(img[..., :3].astype(np.float64))
    acc = None
    for k in range(5):
        g = f.copy()
        g.frame_index, g.reset_accum = k, 1 if k == 0 else 0
        acc, _ = sc.render(g, 32, 0, accum=acc)
    assert np.all(acc[..., 3] == 5.0)
    assert np.allclose(acc[..., :3], np.mean(singles, axis=0), rtol=1e-5, atol=1e-7)
    e = 0.5
    acc = None
    for k in range(5):
        g = f.copy()
        g.frame_index, g.reset_accum, g.ema_alpha = k, 1 if k == 0 else 0, e
        acc, _ = sc.render(g, 32, 0, accum=acc)
    want = singles[0]
    for k in range(1, 5):
        want = e * want + (1.0 - e) * singles[k]
    assert np.allclose(acc[..., :3], want, rtol=1e-5, atol=1e-7)
