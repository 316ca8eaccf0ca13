"""Closed-form pins of the oracle's dielectric, metal, light-tree and seed
arithmetic (CPU only).

Each expected value below follows from optics, geometry or probability, not
from the oracle's code:
  * point light through a glass slab: the light NEE shadow ray crosses two
    parallel interfaces and carries (1 - F1) g (1 - F2) g, with Schlick's F
    at the entry angle and at the Snell-refracted angle inside the slab
    (transmittance, PathTrace.slang:5089-5196);
  * reflection off a glass plane: the reflect branch is taken with
    probability F (Schlick), the throughput is unchanged, so an emitter in
    the mirror direction is seen with mean radiance F Le
    (PathTrace.slang:6709-6767);
  * transmission through a glass slab: an emitter behind the slab along the
    incident direction is seen with mean radiance
    Le (1 - F1)(1 - F2) sum_k F2^2k = Le (1 - F1) / (1 + F2);
  * rough metal at normal incidence: r = normalize(reflect + rho (c - n))
    with c cosine-distributed (PathTrace.slang:6683-6708) is c itself at
    rho = 1 and the half-vector of n and c at rho = 1/2, so the chance of
    hitting an emitter cone of half-angle a is sin^2 a and sin^2 2a;
  * light-tree pick probabilities (PathTrace.slang:3929-4017, node
    summaries of LightTree.cpp:63-147): for point lights, and for clusters of
    coincident point lights, the descent's product of ratios equals
    lum_i cos_i / d_i^2 normalised over the lights; a spot light's cone
    factor is the linear ramp between cos_outer - 0.2 and cos_outer;
  * the light-tree NEE estimator is unbiased: a floor under several point
    lights receives the sum of their closed-form irradiances;
  * the per-pixel seed pcgHash(x*1973 + y*9277 + frame*26699)
    (PathTrace.slang:5457): every pixel of a 7x5 frame draws the jitter the
    published PCG hash predicts (swapped or wrong coefficients fail).
"""
import math

import numpy as np

import oracle_ctypes as orc
import scenes

R0_GLASS = ((1.0 - 1.5) / (1.0 + 1.5)) ** 2


def schlick(cos_t, r0=R0_GLASS):
    return r0 + (1.0 - r0) * (1.0 - cos_t) ** 5


def snell_cos(cos_i, eta):
    """cos of the refracted angle for a ray entering a medium eta times denser."""
    s = math.sqrt(max(0.0, 1.0 - cos_i * cos_i)) / eta
    return math.sqrt(1.0 - s * s)


def _mean_radiance(sd, eye, target, fov=0.02, w=16, h=16, spp=32, frames=40, bounces=4, precision=64, rr=True):
    f = scenes.frame(eye, target, fov, w, h, spp=spp, bounces=bounces)
    f.flags = 0 if rr else 1
    sc = orc.OracleScene.from_scene_data(sd)
    acc = None
    for k in range(frames):
        f.frame_index = k
        f.reset_accum = 1 if k == 0 else 0
        acc, _ = sc.render(f, precision, 0, accum=acc)
    return acc[..., :3].reshape(-1, 3).mean(axis=0)


def test_point_light_through_a_glass_slab_is_exact():
    """Lambert floor, point light above a tinted glass slab (entry face
    normal down at y = 1.5, exit face normal up at y = 1.8).  NEE to a point
    light draws no random number, so one bounce is deterministic:
    a/pi I cos/d^2 (1 - F1) g (1 - F2) g."""
    a = np.array([0.6, 0.5, 0.4])
    g = np.array([0.9, 0.8, 0.7])
    I = np.array([40.0, 40.0, 40.0])
    theta = math.radians(65.0)
    lp = 6.0 * np.array([math.sin(theta), math.cos(theta), 0.0])
    entry = scenes.plane((0, -1, 0), 1.5, albedo=tuple(g), mat=scenes.MAT_DIELECTRIC, ior=1.5)
    exit_ = scenes.plane((0, 1, 0), -1.8, albedo=tuple(g), mat=scenes.MAT_DIELECTRIC, ior=1.5)
    floor = scenes.plane((0, 1, 0), 0.0, albedo=tuple(a))
    sd = scenes.scene([floor, entry, exit_], [scenes.point_light(lp, I)], env=scenes.black_env())
    got = _mean_radiance(sd, (0.0, 1.0, 1.5), (0.0, 0.0, 0.0), w=4, h=4, spp=1, frames=1, bounces=1)
    d = float(np.linalg.norm(lp))
    c1 = lp[1] / d
    c2 = snell_cos(c1, 1.5)
    tr = (1.0 - schlick(c1)) * g * (1.0 - schlick(c2)) * g
    expect = a / math.pi * I * c1 / d ** 2 * tr
    assert np.allclose(got, expect, rtol=3e-4), (got, expect)
    # the refracted angle matters: without refraction the exit Fresnel term
    # would be schlick(c1), a 12 % darker result that this tolerance rejects
    wrong = a / math.pi * I * c1 / d ** 2 * ((1.0 - schlick(c1)) * g) ** 2
    assert not np.allclose(got, wrong, rtol=1e-2)


def test_glass_reflects_with_schlick_probability():
    """A camera ray meets a glass plane at 70 degrees; an emitter fills the
    mirror direction and nothing else is lit: mean radiance = F(70) Le."""
    Le = 2.0
    theta = math.radians(70.0)
    d_in = np.array([math.sin(theta), -math.cos(theta), 0.0])
    mirror_dir = np.array([math.sin(theta), math.cos(theta), 0.0])
    glass = scenes.plane((0, 1, 0), 0.0, albedo=(1, 1, 1), mat=scenes.MAT_DIELECTRIC, ior=1.5)
    emitter = scenes.sphere(tuple(mirror_dir * 10.0), 3.0, albedo=(0, 0, 0), emission=(Le, Le, Le))
    sd = scenes.scene([glass, emitter], [], env=scenes.black_env())
    got = _mean_radiance(sd, tuple(-d_in * 4.0), (0.0, 0.0, 0.0), bounces=3)
    F = schlick(math.cos(theta))
    assert np.allclose(got, F * Le, rtol=0.02), (got, F * Le)
    # a (1 - cos)^4 lobe or an index of 1.33 would be off by > 10 %
    assert abs(F - (R0_GLASS + (1 - R0_GLASS) * (1 - math.cos(theta)) ** 4)) > 0.1 * F


def test_glass_slab_transmits_the_closed_form_series():
    """An emitter behind a clear slab (top face normal up at y = 0, bottom
    face normal down at y = -0.5) along the incident direction at 75
    degrees: the mean radiance is Le (1 - F1)(1 - F2) / (1 - F2^2), with F2
    at the Snell angle inside the slab (internal reflections add the
    geometric series; each round trip costs two bounces)."""
    Le = 2.0
    theta = math.radians(75.0)
    d_in = np.array([math.sin(theta), -math.cos(theta), 0.0])
    top = scenes.plane((0, 1, 0), 0.0, albedo=(1, 1, 1), mat=scenes.MAT_DIELECTRIC, ior=1.5)
    bottom = scenes.plane((0, -1, 0), -0.5, albedo=(1, 1, 1), mat=scenes.MAT_DIELECTRIC, ior=1.5)
    emitter = scenes.sphere(tuple(d_in * 14.0), 6.0, albedo=(0, 0, 0), emission=(Le, Le, Le))
    sd = scenes.scene([top, bottom, emitter], [], env=scenes.black_env())
    got = _mean_radiance(sd, tuple(-d_in * 4.0), (0.0, 0.0, 0.0), bounces=13, rr=False)
    c1 = math.cos(theta)
    c2 = snell_cos(c1, 1.5)
    F1, F2 = schlick(c1), schlick(c2)
    expect = Le * (1.0 - F1) * (1.0 - F2) / (1.0 - F2 * F2)
    assert np.allclose(got, expect, rtol=0.01), (got, expect)
    no_snell = Le * (1.0 - F1) / (1.0 + F1)   # F2 taken at the entry angle
    assert abs(no_snell - expect) > 0.1 * expect


def test_rough_metal_lobe_at_normal_incidence():
    """Metal plane z = 0 seen head on, emitter sphere (radius 4) centred 9
    above it on the normal: from the hit point it fills a cone of half-angle
    a = asin(4/9).  Mean radiance = albedo Le P(reflection in the cone):
    sin^2 a at roughness 1 (cosine lobe), sin^2 2a at roughness 1/2."""
    albedo, Le = 0.5, 2.0
    a = math.asin(4.0 / 9.0)
    for rough, p in ((1.0, math.sin(a) ** 2), (0.5, math.sin(2.0 * a) ** 2)):
        metal = scenes.plane((0, 0, 1), 0.0, albedo=(albedo,) * 3, mat=scenes.MAT_METAL, rough=rough)
        emitter = scenes.sphere((0, 0, 9), 4.0, albedo=(0, 0, 0), emission=(Le, Le, Le))
        sd = scenes.scene([metal, emitter], [], env=scenes.black_env())
        got = _mean_radiance(sd, (0.0, 0.0, 3.0), (0.0, 0.0, 0.0), bounces=3, rr=False)
        expect = albedo * Le * p
        assert np.allclose(got, expect, rtol=0.02), (rough, got, expect)


def _lum(I):
    I = np.asarray(I, np.float64)
    return 0.2126 * I[0] + 0.7152 * I[1] + 0.0722 * I[2]


def _flat_pmf(p, n, lights):
    """lum_i cos_i / d_i^2 normalised (receiver term max(0, cos))."""
    w = []
    for pos, I in lights:
        v = np.asarray(pos, np.float64) - p
        d2 = float(v @ v)
        w.append(_lum(I) * max(0.0, float(n @ v) / math.sqrt(d2)) / d2)
    w = np.array(w)
    return w / w.sum()


def test_light_tree_pmf_of_two_point_lights():
    p, n = np.zeros(3), np.array([0.0, 1.0, 0.0])
    lights = [((1.0, 2.0, 0.0), (3.0, 3.0, 3.0)), ((-4.0, 2.0, 3.0), (6.0, 3.0, 12.0))]
    sd = scenes.scene([scenes.plane((0, 1, 0), 0.0)], [scenes.point_light(*L) for L in lights])
    pmf = orc.OracleScene.from_scene_data(sd).light_pmf(p, n, 2)
    assert np.allclose(pmf, _flat_pmf(p, n, lights), rtol=1e-5), pmf
    # 1/d instead of 1/d^2, or no receiver cosine, would move the split by > 5 %
    v = [np.asarray(L[0]) for L in lights]
    w_d1 = np.array([_lum(L[1]) * (n @ vi) / np.linalg.norm(vi) ** 2 for L, vi in zip(lights, v)])
    assert abs(w_d1[0] / w_d1.sum() - pmf[0]) > 0.05
    # a light below the receiver's horizon keeps only the 1e-8 importance floor
    lights_b = [lights[0], ((0.0, -1.0, 0.0), (5.0, 5.0, 5.0))]
    sd = scenes.scene([scenes.plane((0, 1, 0), 0.0)], [scenes.point_light(*L) for L in lights_b])
    pmf = orc.OracleScene.from_scene_data(sd).light_pmf(p, n, 2)
    imp0 = _lum(lights_b[0][1]) * (n @ np.asarray(lights_b[0][0])) / np.linalg.norm(lights_b[0][0]) ** 3
    assert abs(pmf[1] - 1e-8 / imp0) < 1e-9 and abs(pmf[0] + pmf[1] - 1.0) < 1e-6


def test_light_tree_pmf_of_coincident_clusters():
    """Clusters of coincident point lights: a cluster node's box is the
    lights' shared 1 mm box and its intensity their summed luminance, so the
    two-level descent reduces to the flat lum cos / d^2 distribution.
    Three lights (the lone one splits off first) and four (two pairs)."""
    p, n = np.array([0.2, 0.0, -0.1]), np.array([0.0, 1.0, 0.0])
    A, B = (3.0, 2.0, 0.0), (-1.0, 3.0, 0.5)
    cases = [[(B, (2.0, 2.0, 2.0)), (A, (1.0, 3.0, 0.5)), (A, (4.0, 1.0, 1.0))],
             [(B, (2.0, 2.0, 2.0)), (B, (0.5, 0.5, 3.0)), (A, (1.0, 3.0, 0.5)), (A, (4.0, 1.0, 1.0))]]
    for lights in cases:
        tree = orc.light_tree(np.array([scenes.point_light(*L) for L in lights]))
        assert tree.shape[0] == 2 * len(lights) - 1
        sd = scenes.scene([scenes.plane((0, 1, 0), 0.0)], [scenes.point_light(*L) for L in lights])
        pmf = orc.OracleScene.from_scene_data(sd).light_pmf(p, n, len(lights))
        assert np.allclose(pmf, _flat_pmf(p, n, lights), rtol=1e-5), (pmf, _flat_pmf(p, n, lights))


def test_light_tree_spot_cone_ramp():
    """A spot light (cos_outer = 0.9) beside a point light: the spot's
    importance carries saturate((cos_emit - (cos_outer - 0.2)) / 0.2), with
    cos_emit the cosine between its axis and the direction to the receiver:
    1 at cos_emit = 0.95, 1/2 at 0.8, the 1e-8 floor at 0.6."""
    p, n = np.zeros(3), np.array([0.0, 1.0, 0.0])
    sp, Is = np.array([0.0, 3.0, 0.0]), (2.0, 2.0, 2.0)
    pp, Ip = (2.0, 2.0, 1.0), (1.0, 1.0, 1.0)
    down = -sp / np.linalg.norm(sp)
    for ce, ramp in ((0.95, 1.0), (0.8, 0.5), (0.6, 0.0)):
        # tilt the axis away from the receiver direction by acos(ce)
        t = math.acos(ce)
        axis = np.array([math.sin(t), down[1] * math.cos(t), 0.0])
        spot = scenes.light(1, sp, Is, dir=axis, cos_outer=0.9, cos_inner=0.95)
        sd = scenes.scene([scenes.plane((0, 1, 0), 0.0)], [spot, scenes.point_light(pp, Ip)])
        pmf = orc.OracleScene.from_scene_data(sd).light_pmf(p, n, 2)
        w_s = max(_lum(Is) * ramp * 1.0 / 9.0, 1e-8)
        v = np.asarray(pp) - p
        w_p = _lum(Ip) * (n @ v) / np.linalg.norm(v) ** 3
        assert abs(pmf[0] - w_s / (w_s + w_p)) < 2e-6 * max(1.0, w_s / (w_s + w_p)) + 1e-7, (ce, pmf)


def test_light_tree_nee_sums_the_irradiance_of_many_lights():
    """A Lambert floor under seven point lights picked through the light tree:
    the mean radiance is albedo/pi times the sum of I cos / d^2."""
    a = np.array([0.7, 0.6, 0.5])
    rng = np.random.default_rng(5)
    lights = []
    for _ in range(7):
        pos = rng.uniform([-3, 1, -3], [3, 4, 3])
        I = rng.uniform(2.0, 8.0, 3)
        lights.append((pos, I))
    sd = scenes.scene([scenes.plane((0, 1, 0), 0.0, albedo=tuple(a))], [scenes.point_light(*L) for L in lights],
                      env=scenes.black_env())
    x = np.array([0.3, 0.0, -0.2])
    got = _mean_radiance(sd, x + np.array([0.0, 1.0, 1.5]), x, w=8, h=8, bounces=1)
    E = np.zeros(3)
    for pos, I in lights:
        v = pos - x
        E += I * v[1] / np.linalg.norm(v) ** 3
    assert np.allclose(got, a / math.pi * E, rtol=0.01), (got, a / math.pi * E)


def _pcg(x):
    """Published JCGT 2020 pcg_hash, uint64 arithmetic masked to 32 bits."""
    m = 0xFFFFFFFF
    state = (x * 747796405 + 2891336453) & m
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & m
    return ((word >> 22) ^ word) & m


def test_pixel_seed_on_a_ragged_frame():
    """Every pixel of a 7x5 frame at two frame indices: a one-bounce camera
    ray into the gradient sky (PathTrace.slang:3373-3393) reveals its
    jitter, which must be the two draws after pcgHash(x*1973 + y*9277 +
    frame*26699); swapping the x/y coefficients changes 34 of 35 pixels."""
    sd = scenes.scene(env=None)
    W, H = 7, 5
    for fidx in (0, 5):
        f = scenes.frame((0, 0, 0), (0, 0, -1), 90.0, W, H, bounces=1, frame_index=fidx)
        img, _ = orc.OracleScene.from_scene_data(sd).render(f, 64, 1)
        fwd = np.array(f.fwd_aspect[:3], np.float64)
        right = np.array(f.right_xyz[:3], np.float64)
        up = np.array(f.up_xyz[:3], np.float64)
        aspect, tanf = f.fwd_aspect[3], f.pos_fovtan[3]

        def sky(x, y, cx, cy):
            s1 = _pcg(_pcg((x * cx + y * cy + fidx * 26699) & 0xFFFFFFFF))
            s2 = _pcg(s1)
            jx = (s1 & 0xFFFFFF) / 16777216.0 - 0.5
            jy = (s2 & 0xFFFFFF) / 16777216.0 - 0.5
            ux = ((x + 0.5 + jx) / W) * 2.0 - 1.0
            uy = -(((y + 0.5 + jy) / H) * 2.0 - 1.0)
            d = fwd + right * (ux * aspect * tanf) + up * (uy * tanf)
            d /= np.linalg.norm(d)
            t = min(max(0.5 * (d[1] + 1.0), 0.0), 1.0)
            return np.array([0.5, 0.55, 0.62]) + (1.0 - np.array([0.5, 0.55, 0.62])) * t

        swapped = 0
        for y in range(H):
            for x in range(W):
                assert np.allclose(img[y, x, :3], sky(x, y, 1973, 9277), rtol=1e-6), (fidx, x, y)
                swapped += not np.allclose(img[y, x, :3], sky(x, y, 9277, 1973), rtol=1e-6)
        assert swapped >= 30
