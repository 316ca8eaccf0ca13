"""Closed-form pins of the oracle's sky lights (CPU only): the sun and moon of
the gradient sky, the HDRI cluster lights of an env map, the denoiser's fixed
jitter, and the engine's cluster extraction the C ABI offers beside them.

Expected values come from geometry, probability and the reference's stated
model, computed here in float64:
  * Cephes log within 1 ulp of libm; pow = exp(y log x) within 1e-5;
  * sampleSunDisc (PathTraceMath.slang:55-67): directions inside the 0.00465
    rad disc, uniform over it (P(angle < qR) = q^2, azimuth uniform);
  * sunDisc (PathTrace.slang:3112-3192): at the centre the disc plus the
    full halo, col (250 + 4) atm; past the disc radius the 250 atm term goes;
    halo ((cos - cos_halo) / (1 - cos_halo))^falloff; the oval test at
    vertical flattening 0.5; nothing below y = -0.1;
  * moonDisc (:3212-3353): the dark side of a new moon (halo 0.6 x 0.05),
    the blood moon at exact opposition (eclipse colour), a quarter moon lit
    at the 0.05 floor, the halo-only ring scaled by the phase;
  * the sun NEE of a Lambert floor (:6552-6620) is deterministic: albedo/pi
    sun_rad cos(theta) plus the ambient albedo x skyColorBase(up) = albedo;
    the moon NEE likewise with moon_rad; a sphere between floor and sun
    removes the sun term only; a moon lit under 1 % adds nothing;
  * the HDRI cluster NEE (:6493-6518, sampleHdriLight :3749-3768): with pick
    probabilities proportional to flux x cos the estimator has no variance and
    returns albedo/pi sum(flux_i cos_i);
  * extraction + masked CDF + cluster NEE conserve the env map's irradiance:
    a floor under an env map with a bright sun patch receives albedo/pi times
    the quadrature of L cos over the env map's texels; without the cluster
    lights the patch's share is missing;
  * the fixed jitter (DMT_FRAME_FIXED_JITTER, :5730-5731) puts sample 0
    through (x + 0.5 + jx, y + 0.5 + jy) and draws nothing, so sample 1 takes
    the first two draws of the pixel's seed;
  * the engine's extraction (Engine.cpp:4369-4532): components of the bright
    pixels (columns wrap), < 3-pixel ones dropped, at most 8 kept by
    luminance integral, pick probabilities, centroids, fluxes and the mask;
    the masked CDF is flat on the clusters' texels.
"""
import math

import numpy as np

import dmt
import oracle_ctypes as orc
import scenes
from test_oracle_pins import _pcg

SUN_R = 0.00465
GRAD_LO = np.array([0.5, 0.55, 0.62])


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def _f32(v):
    return np.asarray(v, np.float32).astype(np.float64)


def _base(y):
    """skyColorBase: the gradient from horizon grey-blue to white."""
    t = min(max(0.5 * (y + 1.0), 0.0), 1.0)
    return GRAD_LO + (1.0 - GRAD_LO) * t


def _air(y, warm_lo, warm_hi):
    """(colour, transmittance) of the air mass 1 / max(y, 0.04)."""
    am = 1.0 / max(y, 0.04)
    atm = math.exp(-0.30 * am)
    w = min(max((am - 1.0) * 0.18, 0.0), 1.0)
    return np.array(warm_lo) + (np.array(warm_hi) - np.array(warm_lo)) * w, atm


def _frame(sun=(0, 0, 0), sun_extra=(1, 1, 1), bright=1.0, moon=(0, 0, 0, 0), moon_extra=(1, 1, 1)):
    f = scenes.frame((0, 0, 0), (0, 0, -1), 40.0, 8, 8)
    for i in range(3):
        f.sun[i] = float(sun[i])
        f.sun_extra[i] = float(sun_extra[i])
        f.moon_extra[i] = float(moon_extra[i])
    for i in range(4):
        f.moon[i] = float(moon[i])
    f.sun_disc_brightness = bright
    return f


def _sky(sc, f, dirs):
    return sc.sky(f, np.array(dirs, np.float32)).astype(np.float64)


def test_log_and_pow_against_libm():
    x = np.exp(np.random.default_rng(1).uniform(np.log(1e-30), np.log(1e30), 100000)).astype(np.float32)
    got = orc.math_f32("log", x).astype(np.float64)
    ref = np.log(x.astype(np.float64))
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(got - ref) / ulp) <= 1.0
    t = np.random.default_rng(2).uniform(1e-6, 1.0, 100000).astype(np.float32)
    y = np.random.default_rng(3).uniform(3.0, 8.0, 100000).astype(np.float32)
    got = orc.math_f32("pow", t, y).astype(np.float64)
    ref = np.power(t.astype(np.float64), y.astype(np.float64))
    m = ref > 1e-30
    assert np.max(np.abs(got[m] - ref[m]) / ref[m]) <= 1e-5
    assert orc.math_f32("pow", np.array([1.0, 0.0], np.float32), np.array([7.0, 4.0], np.float32)).tolist() == [1.0, 0.0]


def test_sun_disc_sample_is_uniform_over_the_disc():
    d = _f32(_unit((0.3, 0.8, -0.5)))
    s = orc.sample_sun_disc(d.astype(np.float32), 12345, 40000).astype(np.float64)
    # the angle from the centre, from the cross product (well conditioned at 5 mrad)
    ang = np.arcsin(np.linalg.norm(np.cross(s, d), axis=1) / np.linalg.norm(s, axis=1))
    assert ang.max() <= SUN_R * (1 + 1e-3)
    q = np.sort((ang / SUN_R) ** 2)   # uniform on [0, 1] for an area-uniform disc
    assert np.max(np.abs(q - (np.arange(q.size) + 0.5) / q.size)) < 0.02
    t = _unit(np.cross((0, 1, 0), d))
    b = np.cross(d, t)
    az = np.arctan2(s @ b, s @ t)
    h, _ = np.histogram(az, bins=16, range=(-np.pi, np.pi))
    assert h.min() > 0.9 * s.shape[0] / 16


def test_sun_disc_centre_edge_halo_and_oval():
    sun = _f32(_unit((0.2, 0.6, -0.77)))
    col, atm = _air(sun[1], (1.00, 0.98, 0.94), (1.55, 0.55, 0.20))
    sc = orc.OracleScene.from_scene_data(scenes.scene())
    hz = min(max(1.0 - sun[1], 0.0), 1.0)
    halo_deg = 2.0 + hz * 6.0
    falloff = 8.0 - 5.0 * hz
    cos_halo = math.cos(math.radians(halo_deg))
    r = math.radians(0.55)
    t = _unit(np.cross((0, 1, 0), sun))
    u = np.cross(sun, t)
    dirs = [sun, _f32(_unit(sun + t * math.tan(0.9 * r))), _f32(_unit(sun + t * math.tan(1.1 * r))),
            _f32(_unit(sun + u * math.tan(math.radians(0.5 * halo_deg))))]

    def halo(dd):
        return ((float(np.dot(dd, sun)) - cos_halo) / (1.0 - cos_halo)) ** falloff

    got = _sky(sc, _frame(sun=sun, bright=0.8), dirs)
    want = [col * (250 * atm + 4 * atm) * 0.8,
            (col * 250 * atm + col * halo(dirs[1]) * atm * 4) * 0.8,
            col * halo(dirs[2]) * atm * 4 * 0.8,
            col * halo(dirs[3]) * atm * 4 * 0.8]
    for g, w, d in zip(got, want, dirs):
        # float32 cos near 1 over (1 - cos_halo) ~ 3e-3, to the power ~6: ~1e-3
        assert np.allclose(g - _base(d[1]), w, rtol=2e-3), (g - _base(d[1]), w)
    assert got[1][0] - got[2][0] > 100.0   # the disc edge
    # oval: at vertical flattening 0.5 a horizontal offset of 0.75 r is inside
    # the disc and the same vertical offset outside
    lr = _unit(np.cross((0, 1, 0), sun))
    lu = np.cross(sun, lr)
    dh = _f32(_unit(sun + lr * math.tan(0.75 * r)))
    dv = _f32(_unit(sun + lu * math.tan(0.75 * r)))
    o = _sky(sc, _frame(sun=sun, sun_extra=(1.0, 1.0, 0.5)), [dh, dv])
    assert o[0][0] - _base(dh[1])[0] > 200 * atm and o[1][0] - _base(dv[1])[0] < 10 * atm
    # below y = -0.1 the sun adds nothing
    low = _f32(_unit((0.0, -0.2, -1.0)))
    assert np.allclose(_sky(sc, _frame(sun=low), [low])[0], _base(low[1]), rtol=1e-6)


def test_moon_disc_phases_and_eclipse():
    m = _f32(_unit((-0.3, 0.5, -0.81)))
    sd = scenes.scene()
    sd.moon = np.full((16, 32, 4), 0.8, np.float32)   # a uniform moon surface
    sc = orc.OracleScene.from_scene_data(sd)
    col_atm, m_atm = _air(m[1], (1.0, 1.0, 1.0), (1.20, 0.65, 0.40))
    tc = np.array([0.92, 0.94, 1.00]) * col_atm

    def disc(sun, phase, bright=1.0, d=m):
        # the sun disc is switched off (brightness 0): only the moon adds
        f = _frame(sun=sun, bright=0.0, moon=tuple(m) + (phase,), moon_extra=(1.0, 1.0, bright))
        return _sky(sc, f, [d])[0] - _base(d[1])

    # new moon: the sun in the moon's direction lights the far side; the near
    # side gets the 0.05 floor of the halo (0.6 at the centre)
    assert np.allclose(disc(m, 0.0), tc * 0.6 * m_atm * 0.05, rtol=1e-3)
    # exact opposition: total eclipse, the blood-moon colour (limb 1 at the centre)
    want = np.array([0.55, 0.15, 0.04]) * 0.8 * (1.0 * m_atm * 1.2)
    assert np.allclose(disc(-m, math.pi, bright=0.7), want, rtol=1e-3)
    # quarter moon (sun at right angles): the centre sits on the terminator,
    # lit at the 0.05 floor, no eclipse
    q = _f32(_unit(np.cross(m, (0.0, 1.0, 0.0))))
    want = tc * 0.8 * (3.5 * 0.7 * 0.05 * 1.0 * m_atm) + tc * 0.6 * 0.05 * m_atm * (0.6 * 0.7)
    assert np.allclose(disc(q, math.pi / 2, bright=0.7), want, rtol=1e-3)
    # between the disc (0.55 deg) and the halo (1.5 deg): the halo ring, scaled by the phase
    t = _unit(np.cross((0, 1, 0), m))
    d = _f32(_unit(m + t * math.tan(math.radians(1.0))))
    cos_h = math.cos(math.radians(1.5))
    halo = ((float(np.dot(d, m)) - cos_h) / (1.0 - cos_h)) ** 4 * 0.6
    phase = 2.0
    want = tc * halo * (0.5 * (1.0 - math.cos(phase))) * m_atm * 0.3
    assert np.allclose(disc(q, phase, d=d), want, rtol=5e-3)


def _floor_probe(sd, f, spp=1, frames=1):
    f.spp = spp
    sc = orc.OracleScene.from_scene_data(sd)
    acc = None
    for k in range(frames):
        f.frame_index, f.reset_accum = k, 1 if k == 0 else 0
        acc, _ = sc.render(f, 64, 0, accum=acc)
    return acc[..., :3].reshape(-1, 3).mean(0)


def _floor_frame(sun=None, moon=None):
    """One bounce of a 4x4 frame looking straight at the floor point below (0, 1, 1.5)."""
    f = scenes.frame((0.0, 1.0, 1.5), (0.0, 0.0, 0.0), 0.02, 4, 4, bounces=1)
    f.firefly_clamp = 0.0
    if sun is not None:
        for i in range(3):
            f.sun[i] = float(sun[i])
    if moon is not None:
        for i in range(4):
            f.moon[i] = float(moon[i])
    return f


def test_sun_and_moon_nee_of_a_lambert_floor():
    a = np.array([0.6, 0.5, 0.4])
    floor = scenes.plane((0, 1, 0), 0.0, albedo=tuple(a))
    sd = scenes.scene([floor])   # no env map: the gradient sky
    sun = _f32(_unit((0.4, 0.7, 0.2)))
    col, atm = _air(sun[1], (1.00, 0.97, 0.92), (1.55, 0.55, 0.20))
    got = _floor_probe(sd, _floor_frame(sun=sun))
    want = a / math.pi * (col * 80.0 * atm) * sun[1] + a   # + the ambient skyColorBase(up) = 1
    assert np.allclose(got, want, rtol=1e-5), (got, want)
    # a sphere between the floor and the sun takes the sun term away, not the ambient
    blocker = scenes.sphere(tuple(sun * 3.0), 0.5, albedo=(0.5, 0.5, 0.5))
    got = _floor_probe(scenes.scene([floor, blocker]), _floor_frame(sun=sun))
    assert np.allclose(got, a, rtol=1e-6)
    # the moon, 66 % lit: moon_rad = col (0.3 phase_b atm); its NEE takes the
    # cosine of the jittered direction (within 5 mrad)
    moon = _f32(_unit((-0.5, 0.6, 0.3)))
    phase = 1.9
    pb = 0.5 * (1.0 - math.cos(phase))
    mcol, matm = _air(moon[1], (0.92, 0.94, 1.00), (1.20, 0.65, 0.40))
    got = _floor_probe(sd, _floor_frame(moon=tuple(moon) + (phase,)), spp=16)
    want = a / math.pi * (mcol * (0.3 * pb * matm)) * moon[1]
    assert np.allclose(got - a, want, rtol=3e-3), (got - a, want)
    # a moon lit less than 1 % adds nothing
    got = _floor_probe(sd, _floor_frame(moon=tuple(moon) + (0.1,)))
    assert np.allclose(got, a, rtol=1e-6)


def test_hdri_cluster_nee_has_no_variance_with_matched_probabilities():
    a = np.array([0.7, 0.6, 0.5])
    floor = scenes.plane((0, 1, 0), 0.0, albedo=tuple(a))
    sd = scenes.scene([floor], env=np.zeros((8, 16, 3), np.float32))
    d1, d2 = _f32(_unit((0.3, 0.9, 0.1))), _f32(_unit((-0.6, 0.5, 0.4)))
    f1, f2 = np.array([3.0, 2.5, 2.0]), np.array([1.0, 1.2, 0.8])
    w = np.array([f1[1] * d1[1], f2[1] * d2[1]])   # pick ~ flux x cos in green: no variance in green
    p = w / w.sum()
    dirs = [tuple(d1) + (p[0],), tuple(d2) + (p[1],)]
    cols = [tuple(f1) + (0.0,), tuple(f2) + (0.0,)]
    f = _floor_frame()
    dmt.set_hdri_lights(f, 2, dirs, cols)
    got = _floor_probe(sd, f, spp=16, frames=8)
    want = a / math.pi * (f1 * d1[1] + f2 * d2[1])
    assert abs(got[1] - want[1]) < 1e-5 * want[1]
    assert np.allclose(got, want, rtol=0.03)
    # without the env map the cluster lights do nothing (the gradient sky's ambient only)
    f = _floor_frame()
    dmt.set_hdri_lights(f, 2, dirs, cols)
    assert np.allclose(_floor_probe(scenes.scene([floor]), f, spp=4), a, rtol=1e-6)


def _texel_quadrature(env, weight=None):
    """sum over texels of L_green max(cos theta, 0) dOmega (texel centres,
    dOmega = 2 pi^2 / (W H) sin theta), optionally times a 0/1 weight."""
    H, W = env.shape[:2]
    th = (np.arange(H) + 0.5) / H * np.pi
    dw = (2 * np.pi * np.pi / (W * H)) * np.sin(th)[:, None]
    e = env[..., 1].astype(np.float64) * np.maximum(np.cos(th), 0.0)[:, None] * dw
    return float((e if weight is None else e * weight).sum())


def test_cluster_extraction_conserves_the_env_irradiance():
    """Uniform sky 0.2 plus a 3x3 sun patch of 2000 texels: extraction takes
    the patch out of the importance CDF and gives it its own NEE.  At the
    last bounce the env NEE is unweighted and samples texel centres, so the
    floor receives albedo/pi (texel quadrature over the unmasked texels +
    flux x cos of the cluster), which is the quadrature over the whole map."""
    H, W = 64, 128
    env = np.full((H, W, 3), 0.2, np.float32)
    env[10:13, 40:43] = 2000.0
    n, dirs, cols, mask = dmt.extract_env_lights(env)
    assert n == 1 and mask.sum() == 9 and mask[10:13, 40:43].all()
    m, c, total = dmt.build_env_cdf(env, mask)
    a = 0.5
    floor = scenes.plane((0, 1, 0), 0.0, albedo=(a, a, a))
    sd = scenes.scene([floor], env=env)
    sd.env_cdf = (m, c, total)
    q_rest = _texel_quadrature(env, 1 - mask)
    q_all = _texel_quadrature(env)
    cluster = float(cols[0, 1]) * float(dirs[0, 1])
    assert abs(q_rest + cluster - q_all) < 2e-3 * q_all   # the centroid stands for the patch
    f = _floor_frame()
    dmt.set_hdri_lights(f, n, dirs[:n], cols[:n])
    got = _floor_probe(sd, f, spp=16, frames=16)[1]
    assert abs(got - a / math.pi * (q_rest + cluster)) < 3e-3 * got, (got, a / math.pi * (q_rest + cluster))
    # without the cluster lights the masked CDF leaves the patch out
    got0 = _floor_probe(sd, _floor_frame(), spp=16, frames=16)[1]
    assert abs(got0 - a / math.pi * q_rest) < 0.02 * a / math.pi * q_rest
    assert got0 < 0.1 * got


def test_fixed_jitter_replaces_the_first_sample_draws():
    sd = scenes.scene()
    f = scenes.frame((0, 0, 0), (0, 0, -1), 90.0, 1, 1, bounces=1, spp=1)
    f.flags = dmt.DMT_FRAME_FIXED_JITTER
    f.jitter[0], f.jitter[1] = 0.375, -0.25
    sc = orc.OracleScene.from_scene_data(sd)
    fwd = np.array(f.fwd_aspect[:3], np.float64)
    right = np.array(f.right_xyz[:3], np.float64)
    up = np.array(f.up_xyz[:3], np.float64)
    aspect, tanf = f.fwd_aspect[3], f.pos_fovtan[3]

    def sky_at(jx, jy):
        ux = (0.5 + jx) * 2.0 - 1.0
        uy = -((0.5 + jy) * 2.0 - 1.0)
        return _base(_unit(fwd + right * (ux * aspect * tanf) + up * (uy * tanf))[1])

    one, _ = sc.render(f, 64, 1)
    assert np.allclose(one[0, 0, :3], sky_at(0.375, -0.25), rtol=1e-6)
    f.spp = 2
    two, _ = sc.render(f, 64, 1)
    s1 = _pcg(_pcg(0))   # pixel (0, 0) of frame 0: seed pcgHash(0); two draws
    s2 = _pcg(s1)
    jx = (s1 & 0xFFFFFF) / 16777216.0 - 0.5
    jy = (s2 & 0xFFFFFF) / 16777216.0 - 0.5
    assert np.allclose(two[0, 0, :3], 0.5 * (sky_at(0.375, -0.25) + sky_at(jx, jy)), rtol=1e-6)


def _blob(env, v0, u0, h, w, val):
    W = env.shape[1]
    for v in range(v0, v0 + h):
        for u in range(u0, u0 + w):
            env[v, u % W] = val


def test_extraction_components_probabilities_and_mask():
    H, W = 32, 64
    env = np.full((H, W, 3), 0.1, np.float32)
    _blob(env, 4, 10, 3, 3, (60.0, 50.0, 40.0))     # 9 texels
    _blob(env, 20, 62, 2, 4, (30.0, 30.0, 30.0))    # 8 texels across the u wrap
    _blob(env, 12, 30, 1, 2, (80.0, 80.0, 80.0))    # 2 texels: too small, stays in the CDF
    n, dirs, cols, mask = dmt.extract_env_lights(env, top_frac=0.02)
    assert n == 2
    assert mask.sum() == 17 and mask[12, 30] == 0 and mask[20, 63] == 1 and mask[20, 0] == 1
    th = (np.arange(H) + 0.5) / H * np.pi
    ph = (np.arange(W) + 0.5) / W * 2 * np.pi - np.pi
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    dw = (2 * np.pi * np.pi / (W * H)) * np.sin(TH)
    v = np.stack([np.sin(TH) * np.cos(PH), np.cos(TH), np.sin(TH) * np.sin(PH)], -1)
    e = env.astype(np.float64)
    lum = 0.2126 * e[..., 0] + 0.7152 * e[..., 1] + 0.0722 * e[..., 2]
    A = np.zeros((H, W), bool)
    A[4:7, 10:13] = True
    B = np.zeros((H, W), bool)
    B[20:22, 62:64] = True
    B[20:22, 0:2] = True
    comps = []
    for sel in (A, B):
        L = (lum * dw)[sel].sum()
        cen = (v[sel] * (lum * dw)[sel][:, None]).sum(0)
        flux = (e[sel] * dw[sel][:, None]).sum(0)
        comps.append((L, cen / np.linalg.norm(cen), flux))
    comps.sort(key=lambda t: -t[0])
    tot = sum(t[0] for t in comps)
    for i, (L, cen, flux) in enumerate(comps):
        assert np.allclose(dirs[i, :3], cen, atol=1e-6)
        assert np.allclose(cols[i, :3], flux, rtol=1e-6)
        assert abs(dirs[i, 3] - L / tot) < 1e-6
    # the masked CDF does not rise on the clusters' texels (isExtractedClusterDir)
    m, c, total = dmt.build_env_cdf(env, mask)
    prev = np.concatenate([np.zeros((H, 1), np.float32), c[:, :-1]], 1)
    assert np.all((c <= prev)[mask == 1]) and np.all((c > prev)[mask == 0])
    # at most eight clusters, the brightest by luminance integral
    env2 = np.full((H, W, 3), 0.1, np.float32)
    for k in range(10):
        _blob(env2, 2 + 3 * (k // 5), 2 + 12 * (k % 5), 2, 2, (10.0 + 5.0 * k,) * 3)
    n2, d2, c2, m2 = dmt.extract_env_lights(env2, top_frac=0.05)
    assert n2 == 8 and m2.sum() == 32
    assert m2[2:4, 2:4].sum() == 0 and m2[2:4, 14:16].sum() == 0   # the two dimmest stay in the CDF
    assert d2[0, 3] > d2[7, 3] > 0 and abs(d2[:8, 3].sum() - 1.0) < 1e-6
    # a dim image has no lights
    assert dmt.extract_env_lights(np.full((8, 16, 3), 0.3, np.float32))[0] == 0
