"""Pins the fp64 tangent-linear oracle (orc_tangent): its derivative of the
frame value along one material parameter must equal a central finite
difference of the fp64 oracle's own image, parameter by parameter, with a step
small enough that no sample's decision flips (albedo and emission move no
ray: h = 1e-3; roughness bends rays: h = 1e-4)."""
import numpy as np

import dmt
import oracle_ctypes as orc


def _fd(sd, f, prim, k, dimg, h=1e-3):
    col = [4, 5, 6, 7, 16, 17, 18][k]
    p0 = np.float32(sd.prims[prim, col])
    vals = []
    steps = []
    for sgn in (1.0, -1.0):
        P = sd.prims.copy()
        P[prim, col] = np.float32(p0 + sgn * h)
        steps.append(float(P[prim, col]) - float(p0))
        o = orc.OracleScene(P, sd.lights, sd.positions, sd.indices, sd.env, sd.env_intensity)
        vals.append(float((dimg[..., :3].astype(np.float64) * o.render_f64(f)).sum()))
    return (vals[0] - vals[1]) / (steps[0] - steps[1])


def test_tangent_matches_fp64_finite_differences():
    sd = dmt.generate(4, seed=5, spheres=24, width=48, height=32)
    f = sd.frame.copy()
    f.spp = 2
    f.flags = dmt.DMT_FRAME_NO_RUSSIAN_ROULETTE
    osc = orc.OracleScene.from_scene_data(sd)
    dimg = np.random.default_rng(0).normal(size=(32, 48, 4)).astype(np.float32)
    mats = sd.prims[:, 9].astype(int)
    checked = 0
    for prim in range(1, 13):
        for k in range(7):
            if k == 3 and mats[prim] != 1:
                assert osc.tangent(f, prim, k, dimg) == 0.0   # roughness acts on metal only
                continue
            if (prim + k) % 3:
                continue
            g = osc.tangent(f, prim, k, dimg)
            if k == 3:
                # roughness moves rays: a larger step lets a few samples cross
                # a visibility edge, which the interior derivative leaves out
                fd = _fd(sd, f, prim, k, dimg, h=1e-4)
                assert abs(g - fd) <= 1e-3 * abs(fd) + 1e-7, (prim, k, g, fd)
            else:
                fd = _fd(sd, f, prim, k, dimg)
                assert abs(g - fd) <= 1e-6 * abs(fd) + 1e-9, (prim, k, g, fd)
            checked += 1
    assert checked >= 15


def test_tangent_is_linear_in_the_image_gradient():
    sd = dmt.generate(4, seed=6, spheres=12, width=24, height=16)
    f = sd.frame.copy()
    osc = orc.OracleScene.from_scene_data(sd)
    rng = np.random.default_rng(1)
    a = rng.normal(size=(16, 24, 4)).astype(np.float32)
    b = rng.normal(size=(16, 24, 4)).astype(np.float32)
    ga, gb = osc.tangent(f, 3, 1, a), osc.tangent(f, 3, 1, b)
    gab = osc.tangent(f, 3, 1, a + b)
    assert abs(gab - (ga + gb)) <= 1e-5 * (abs(ga) + abs(gb)) + 1e-9
    _, img = osc.tangent(f, 3, 5, a, want_image=True)
    # dI/d(emission.g) is zero in the red and blue channels
    assert np.all(img[..., 0] == 0.0) and np.all(img[..., 2] == 0.0) and np.any(img[..., 1] != 0.0)
