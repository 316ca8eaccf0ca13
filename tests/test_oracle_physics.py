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
    (PathTrace.slang:6709-6767), so every escaping path carries L=1 exactly.
Wrong pdfs (area->solid angle), a dropped 1/pi, a wrong MIS weight or a wrong
Fresnel branch each fail one of them.
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
