"""Pins of the oracle's camera-sample features (PathTrace.slang:5729-5816,
PathTraceMath.slang:96-310): the Cephes exp against libm; the spectral weight
against the CIE 1931 observer and the band-mean normaliser the reference
states; Cauchy's law against its d-line anchor and the crown-glass Abbe number;
the lens samplers against their closed forms and densities; the shutter-time
trace against closed-form hit distances and brute force; and image-level
identities of motion blur."""
import math

import numpy as np
import pytest

import oracle_ctypes as orc
from scenes import MAT_DIELECTRIC, frame, quad_light, scene, sphere

# CIE 1931 2-degree observer (x-bar, y-bar, z-bar), tabulated values.
CIE1931 = {450: (0.3362, 0.0380, 1.7721), 500: (0.0049, 0.3230, 0.2720), 555: (0.5121, 1.0000, 0.0057),
           600: (1.0622, 0.6310, 0.0008), 650: (0.2835, 0.1070, 0.0000)}
# linear sRGB (D65) -> XYZ, the inverse of the matrix the reference applies
SRGB_TO_XYZ = np.array([[0.4124564, 0.3575761, 0.1804375], [0.2126729, 0.7151522, 0.0721750],
                        [0.0193339, 0.1191920, 0.9503041]])
INV_BAND_MEAN = np.array([2.730792, 3.443936, 3.606855])   # PathTraceMath.slang:190


def _weight(lam):
    lam = np.atleast_1d(np.asarray(lam, np.float32))
    return np.stack([orc.math_f32(f"spectral_{c}", lam) for c in "rgb"], 1)


def test_exp_matches_libm():
    x = np.linspace(-86.9, 88.0, 400001).astype(np.float32)
    got = orc.math_f32("exp", x).astype(np.float64)
    ref = np.exp(x.astype(np.float64))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(got - ref) / ulp) <= 1.0
    edge = orc.math_f32("exp", np.array([-87.5, -200.0, 0.0, 89.0], np.float32))
    assert edge[0] == 0.0 and edge[1] == 0.0 and edge[2] == 1.0 and np.isinf(edge[3])


def test_spectral_weight_has_unit_band_mean():
    """spectralWeightSRGB is normalised so E[weight] over a uniform wavelength
    in [380, 730] nm is (1, 1, 1) (PathTraceMath.slang:169-192)."""
    lam = np.arange(380.5, 730.0, 1.0)
    assert np.allclose(_weight(lam).mean(0), 1.0, atol=1e-5)


def test_spectral_weight_reconstructs_the_cie_observer():
    """Undoing the band means and the sRGB matrix gives back the Wyman-Sloan-
    Shirley fit, which must sit on the CIE 1931 table."""
    for lam, ref in CIE1931.items():
        xyz = SRGB_TO_XYZ @ (_weight(lam)[0].astype(np.float64) / INV_BAND_MEAN)
        assert np.allclose(xyz, ref, atol=0.015), (lam, xyz, ref)


def test_cauchy_index_anchors_at_the_d_line():
    base = np.float32(1.5168)
    nd, nf, nc = orc.math_f32("cauchy", np.array([587.6, 486.1, 656.3], np.float32), np.full(3, base, np.float32))
    assert abs(nd - base) <= 2 * np.spacing(base)
    # BK7: C = 0.0042 um^2 gives the crown-glass Abbe number ~64 (PathTraceMath.slang:102-129)
    assert 63.5 < (nd - 1.0) / (nf - nc) < 65.5
    lam = np.linspace(380, 730, 64).astype(np.float32)
    n = orc.math_f32("cauchy", lam, np.full(lam.size, base, np.float32))
    assert np.all(np.diff(n) < 0)   # normal dispersion: blue bends more


def test_concentric_disk_closed_forms_and_density():
    pts = orc.lens_sample(np.array([[1.0, 0.5], [0.0, 0.5], [0.5, 1.0], [0.5, 0.5]], np.float32))
    assert np.allclose(pts, [[1, 0], [-1, 0], [0, 1], [0, 0]], atol=1e-6)
    u = np.random.default_rng(1).random((200000, 2), dtype=np.float32)
    p = orc.lens_sample(u)
    r = np.hypot(p[:, 0], p[:, 1])
    assert r.max() <= 1.0 + 1e-6
    for q in (0.25, 0.5, 0.75):   # area-uniform: P(r < q) = q^2
        assert abs((r < q).mean() - q * q) < 0.005
    assert np.allclose(p.mean(0), 0.0, atol=0.005)


@pytest.mark.parametrize("blades,rot", [(3, 0.0), (5, 0.4), (6, 0.3), (8, 1.0)])
def test_polygon_aperture_is_area_uniform(blades, rot):
    u = np.random.default_rng(blades).random((200000, 2), dtype=np.float32)
    p = orc.lens_sample(u, blades=blades, rot=rot).astype(np.float64)
    # outward edge normals of the regular N-gon of circumradius 1 rotated by rot
    mid = rot + (np.arange(blades) + 0.5) * 2.0 * math.pi / blades
    normals = np.stack([np.cos(mid), np.sin(mid)], 1)
    apothem = math.cos(math.pi / blades)
    depth = (p @ normals.T).max(1) / apothem   # 1 on the boundary
    assert depth.max() <= 1.0 + 1e-5
    for q in (0.5, 0.8):   # the polygon scaled by q holds q^2 of the area
        assert abs((depth < q).mean() - q * q) < 0.005
    wedge = np.floor(((np.arctan2(p[:, 1], p[:, 0]) - rot) % (2 * math.pi)) / (2 * math.pi / blades))
    counts = np.bincount(wedge.astype(int), minlength=blades)
    assert np.all(np.abs(counts / counts.mean() - 1.0) < 0.03)


def test_cat_eye_narrows_the_radial_component():
    ap = np.array([[0.9, 0.7], [0.2, 0.55], [0.1, 0.95]], np.float32)
    plain = orc.lens_sample(ap)
    assert np.array_equal(orc.lens_sample(ap, cat_eye=1.0, uv=np.zeros((3, 2), np.float32)), plain)
    for uv, s in (((1.0, 0.0), 1.0), ((0.6, -0.8), 0.5), ((0.3, 0.4), 0.8), ((2.0, 2.0), 1.0)):
        got = orc.lens_sample(ap, cat_eye=s, uv=np.tile(np.array(uv, np.float32), (3, 1)))
        e = min(math.hypot(*uv), 1.0)
        radial = np.array(uv) / math.hypot(*uv)
        tangent = np.array([-radial[1], radial[0]])
        assert np.allclose(got @ tangent, plain @ tangent, atol=1e-6)
        assert np.allclose(got @ radial, (plain @ radial) * (1.0 - 0.7 * s * e), atol=1e-6)


def _moving_spheres(n=64, seed=4):
    rng = np.random.default_rng(seed)
    prims = []
    for i in range(n):
        c = rng.uniform(-4, 4, 3)
        p = sphere(c, float(rng.uniform(0.2, 0.8)))
        p[12:15] = c - rng.uniform(-1.5, 1.5, 3)   # the previous centre
        prims.append(p)
    return prims


def test_shutter_time_moves_spheres_from_previous_to_current():
    prims = _moving_spheres(24)
    s = sphere((0.5, 0.0, 10.0), 1.0)
    s[12:15] = (-0.5, 0.0, 10.0)
    prims.append(s)
    osc = orc.OracleScene.from_scene_data(scene(prims))
    o = np.array([[0.0, 0.0, 0.0]], np.float32)
    d = np.array([[0.0, 0.0, 1.0]], np.float32)
    for t in (0.0, 0.25, 0.5, 0.9, 1.0):
        cx = -0.5 + t
        h = osc.trace_at(o, d, t)[0]
        assert h["hit"] == 1 and h["prim"] == len(prims) - 1
        assert abs(h["t"] - (10.0 - math.sqrt(1.0 - cx * cx))) < 2e-5


def test_shutter_time_translates_the_mesh():
    pos = np.array([[-5, -5, 10], [5, -5, 10], [0, 5, 10]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    osc = orc.OracleScene.from_scene_data(scene(positions=pos, indices=idx))
    o = np.array([[0.0, 0.0, 0.0]], np.float32)
    d = np.array([[0.0, 0.0, 1.0]], np.float32)
    shift = ((0.0, 0.0, 0.0), (0.0, 0.0, 2.0))
    for t in (0.0, 0.3, 1.0):
        h = osc.trace_at(o, d, t, shift)[0]
        assert h["hit"] == 2 and abs(h["t"] - (10.0 + 2.0 * t)) < 1e-5
    assert abs(osc.trace_at(o, d, 1.0)[0]["t"] - 10.0) < 1e-6


def test_swept_boxes_keep_every_hit():
    """The sphere BVH holds each sphere over its whole motion: BVH and brute
    force agree at any shutter time."""
    osc = orc.OracleScene.from_scene_data(scene(_moving_spheres(96)))
    rng = np.random.default_rng(9)
    n = 4000
    o = rng.uniform(-7, 7, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    for t in (0.0, 0.37, 0.81, 1.0):
        a = osc.trace_at(o, d, t)
        b = osc.trace_at(o, d, t, brute=True)
        assert np.array_equal(a["hit"], b["hit"]) and np.array_equal(a["prim"], b["prim"])
        assert np.array_equal(a["t"], b["t"])
        assert 200 < np.count_nonzero(a["hit"]) < n


def _blur_scene():
    s = sphere((0.8, 0.0, 0.0), 0.9, emission=(4.0, 4.0, 4.0))
    s[12:15] = (-0.8, 0.0, 0.0)
    return scene([s], env=np.zeros((4, 8, 3), np.float32))


def test_motion_blur_off_renders_the_current_positions():
    sd = scene(_moving_spheres(40), [quad_light((0, 6, 0), (0, -1, 0), (1, 0, 0), 1.0, (8, 8, 8))])
    still = scene([np.concatenate([p[:12], p[0:3], p[15:]]) for p in sd.prims], sd.lights)
    f = frame((0.0, 2.0, 12.0), (0.0, 0.0, 0.0), 50.0, 24, 16, spp=2)
    a, _ = orc.OracleScene.from_scene_data(sd).render(f)
    b, _ = orc.OracleScene.from_scene_data(still).render(f)
    assert np.array_equal(a, b)


def test_motion_blur_is_the_shutter_average():
    """An emissive sphere seen once (one bounce): the blurred image converges
    to the mean of still images over the shutter."""
    sd = _blur_scene()
    f = frame((0.0, 0.0, 6.0), (0.0, 0.0, 0.0), 40.0, 24, 12, spp=32, bounces=1)
    f.motion_blur = 1
    osc = orc.OracleScene.from_scene_data(sd)
    acc = None
    for k in range(8):
        g = f.copy()
        g.frame_index, g.reset_accum = k, 1 if k == 0 else 0
        acc, _ = osc.render(g, 32, 0, accum=acc)
    stills = np.zeros((12, 24, 3))
    times = (np.arange(32) + 0.5) / 32.0
    for t in times:
        p = sd.prims[0].copy()
        c = (1.0 - t) * p[12:15] + t * p[0:3]
        p[0:3], p[12:15] = c, c
        g = f.copy()
        g.motion_blur, g.spp = 0, 8
        img, _ = orc.OracleScene.from_scene_data(scene([p], env=sd.env)).render(g)
        stills += img[..., :3] / len(times)
    blurred = acc[..., :3]
    assert abs(blurred.mean() - stills.mean()) < 0.02 * stills.mean()
    # the horizontal profile (sums over rows and channels) along the motion
    prof_b, prof_s = blurred.sum((0, 2)), stills.sum((0, 2))
    assert np.abs(prof_b - prof_s).sum() < 0.05 * prof_s.sum()
    # the streak: columns the sphere covers part of the time are partly lit
    assert np.count_nonzero((blurred[6, :, 0] > 0.2) & (blurred[6, :, 0] < 3.8)) >= 4


def test_spectral_off_and_equal_indices_match_rgb():
    """Without dispersion (C = 0) every wavelength takes the material's index:
    the spectral image is the RGB image reweighted, which averages back to it."""
    sd = scene([sphere((0, 0, 0), 1.5, mat=MAT_DIELECTRIC, ior=1.5, albedo=(1, 1, 1))],
               [quad_light((0, 6, 0), (0, -1, 0), (1, 0, 0), 1.0, (8, 8, 8))],
               env=np.full((8, 16, 3), 0.5, np.float32))
    f = frame((0.0, 1.0, 7.0), (0.0, 0.0, 0.0), 40.0, 16, 12, spp=32)
    rgb, _ = orc.OracleScene.from_scene_data(sd).render(f)
    g = f.copy()
    g.spectral, g.cauchy_c = 1, 0.0
    osc = orc.OracleScene.from_scene_data(sd)
    acc = None
    for k in range(8):
        h = g.copy()
        h.frame_index, h.reset_accum = k, 1 if k == 0 else 0
        acc, _ = osc.render(h, 32, 0, accum=acc)
    assert abs(acc[..., :3].mean() - rgb[..., :3].mean()) < 0.03 * rgb[..., :3].mean()


def test_spectral_accumulation_keeps_signed_values_at_one_spp():
    """At 1 spp a spectral frame value is L * weight(lambda), whose channels
    are negative for out-of-gamut wavelengths; they must reach the running
    mean (PathTrace.slang:7722-7734), so that the mean over frames converges
    to L (the weight's band mean is 1).  Zeroing negatives per frame raises
    the red channel by tens of percent, which fails here."""
    sd = scene(env=np.full((8, 16, 3), 0.5, np.float32))   # camera sees only a uniform grey sky
    f = frame((0.0, 0.0, 5.0), (0.0, 0.0, 0.0), 40.0, 32, 16, spp=1)
    f.spectral, f.cauchy_c = 1, 0.0
    osc = orc.OracleScene.from_scene_data(sd)
    acc = None
    first_negative = False
    nframes = 160
    for k in range(nframes):
        g = f.copy()
        g.frame_index, g.reset_accum = k, 1 if k == 0 else 0
        acc, _ = osc.render(g, 32, 0, accum=acc)
        if k == 0:
            first_negative = bool((acc[..., :3] < 0.0).any())
    assert first_negative, "a single spectral frame at 1 spp must keep negative channels"
    assert np.all(acc[..., 3] == nframes)
    mean = acc[..., :3].reshape(-1, 3).mean(0)
    # 81,920 wavelength samples: the mean of each channel is within ~1 % of 0.5
    assert np.allclose(mean, 0.5, rtol=0.03), mean
