"""Sky lights through the C ABI, against the oracle bit for bit: the gradient
sky's sun and moon (discs seen by the camera, their NEE on the surfaces, a
moon texture), the HDRI cluster lights of an env map with the engine's masked
CDF, the denoiser's fixed jitter; a batch mixing sky frames with a plain one
equal to the per-scene renders; the engine path (PtPush sky fields, the moon
map as an RGBA16F texture at slot 7, the env CDF buffers at slots 4 and 5)
equal to the direct render; the adjoint refusing sky frames."""
import numpy as np
import pytest

import _dmt_rhi as rhi
import dmt
import oracle_ctypes as orc
from diffrender import adjoint_numpy
from rhi_engine import EngineSim

pytestmark = pytest.mark.gpu

RGBA16F, STORAGE = 3, 2


def _unit(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v)).astype(np.float32)


def _both(device, sd, f, frames=2):
    sc = dmt.Scene(device, sd)
    sc.render_frames(f, frames)
    got = sc.readback(f.width, f.height)
    counts = sc.measure_traffic(f)
    sc.close()
    osc = orc.OracleScene.from_scene_data(sd)
    acc = None
    for k in range(frames):
        g = f.copy()
        g.frame_index = f.frame_index + k
        g.reset_accum = 1 if k == 0 else 0
        acc, _ = osc.render(g, 32, 0, accum=acc)
    _, oc = osc.render(f, 32, 0)
    return got, acc, counts, oc


def _check(label, got, ref, counts=None, oc=None):
    same = float(np.mean(np.all(got == ref, axis=-1)))
    print(f"{label}: equal to the oracle on {same:.6f} of pixels", flush=True)
    assert np.array_equal(got, ref), label
    if counts is not None:
        assert (counts["primary_rays"], counts["extension_rays"], counts["shadow_rays"]) == (
            oc["primary"], oc["extension"], oc["shadow"]), label


def _moon_texture(seed=3, w=64, h=32):
    rng = np.random.default_rng(seed)
    t = np.ones((h, w, 4), np.float32)
    t[..., :3] = rng.uniform(0.2, 1.0, (h, w, 3))
    return t.astype(np.float16).astype(np.float32)   # values an RGBA16F texture holds


def _gradient_sky_scene():
    """The sphere workload under the gradient sky (no env map), with a moon
    texture, seen from outside so that the upper half of the frame is sky."""
    s = dmt.generate(0, seed=17, width=48, height=36)
    f = dmt.look_at((0.0, 2.0, 13.0), (0.0, 1.0, 0.0), 50.0, 48, 36, base=s.frame)
    return dmt.SceneData(prims=s.prims, lights=s.lights, positions=s.positions, indices=s.indices, env=None,
                         frame=f, moon=_moon_texture())


def _sun_moon(f, sun_size=2.0, vflat=0.6, bright=0.5, moon_size=3.0, phase=2.2):
    """Sun and moon inside the camera's view, both above the horizon."""
    fwd = np.array(f.fwd_aspect[:3], np.float64)
    up = np.array(f.up_xyz[:3], np.float64)
    right = np.array(f.right_xyz[:3], np.float64)
    sun = _unit(fwd + 0.3 * up + 0.25 * right)
    moon = _unit(fwd + 0.25 * up - 0.35 * right)
    assert sun[1] > 0.05 and moon[1] > 0.05
    f.sun[:] = (*sun, 0.0)
    f.sun_extra[:] = (sun_size, 1.5, vflat, 0.0)
    f.sun_disc_brightness = bright
    f.moon[:] = (*moon, phase)
    f.moon_extra[:] = (moon_size, 1.0, 0.8, 0.0)
    return f


def _hdri_scene(seed=42):
    """The mesh workload whose env map gets three bright patches (one across
    the column wrap): extracted as cluster lights, masked out of the CDF."""
    m = dmt.generate(1, seed=seed, subdivisions=3, random_tris=500, width=40, height=30)
    env = m.env.copy()
    env[60:64, 100:106] = (400.0, 380.0, 300.0)
    env[90:93, 508:512] = (150.0, 160.0, 200.0)
    env[90:93, 0:2] = (150.0, 160.0, 200.0)
    env[40:42, 300:303] = (250.0, 90.0, 40.0)
    n, dirs, cols, mask = dmt.extract_env_lights(env)
    assert n == 3, n
    sd = dmt.SceneData(prims=m.prims, lights=m.lights, positions=m.positions, indices=m.indices, env=env,
                       env_intensity=m.env_intensity, frame=m.frame.copy(), env_cdf=dmt.build_env_cdf(env, mask))
    f = sd.frame.copy()
    dmt.set_hdri_lights(f, n, dirs[:n], cols[:n])
    return sd, f


def test_gradient_sky_sun_and_moon_equal_the_oracle(device):
    sd = _gradient_sky_scene()
    f = _sun_moon(sd.frame.copy())
    f.spp = 2
    got, ref, counts, oc = _both(device, sd, f)
    _check("sun + moon + moon texture", got, ref, counts, oc)
    # the discs and their NEE change the image
    plain, _, _, _ = _both(device, sd, sd.frame.copy())
    assert not np.array_equal(got, plain)


@pytest.mark.parametrize("case", ["new moon", "eclipse", "sun only", "round sun, no texture"])
def test_sky_light_variants_equal_the_oracle(device, case):
    sd = _gradient_sky_scene()
    f = sd.frame.copy()
    if case == "new moon":
        _sun_moon(f, phase=0.05)
        f.moon[:3] = f.sun[:3]
    elif case == "eclipse":
        _sun_moon(f, phase=np.pi)
        f.sun[:3] = tuple(-np.array(f.moon[:3]))
    elif case == "sun only":
        _sun_moon(f)
        f.moon[:] = (0.0, -0.5, 0.866, 1.0)
    else:
        _sun_moon(f, sun_size=1.0, vflat=1.0, bright=1.0)
        sd.moon = None
    got, ref, counts, oc = _both(device, sd, f)
    _check(case, got, ref, counts, oc)


def test_hdri_cluster_lights_equal_the_oracle(device):
    sd, f = _hdri_scene()
    f.spp = 2
    got, ref, counts, oc = _both(device, sd, f)
    _check("HDRI cluster lights", got, ref, counts, oc)
    g = f.copy()
    g.mis_enabled = 0
    got, ref, counts, oc = _both(device, sd, g)
    _check("HDRI cluster lights, MIS off", got, ref, counts, oc)


@pytest.mark.parametrize("spp", [1, 3])
def test_fixed_jitter_equals_the_oracle(device, spp):
    s = dmt.generate(0, seed=21, width=40, height=28)
    f = s.frame.copy()
    f.spp = spp
    f.flags |= dmt.DMT_FRAME_FIXED_JITTER
    f.jitter[0], f.jitter[1] = 0.3125, -0.4375
    got, ref, _, _ = _both(device, s, f)
    _check(f"fixed jitter, {spp} spp", got, ref)


@pytest.mark.parametrize("plain", [0, 1], ids=["graph", "plain launches"])
def test_batch_with_sky_frames_equals_per_scene_renders(device, plain):
    a = _gradient_sky_scene()
    fa = _sun_moon(a.frame.copy())
    b, fb = _hdri_scene(seed=43)
    c = dmt.generate(0, seed=19, width=36, height=24)
    fc = c.frame.copy()
    data = (a, b, c)
    frames = [fa, fb, fc]
    scenes = [dmt.Scene(device, sd) for sd in data]
    batch = dmt.Batch(device, scenes, frames, dmt.BatchParams(plain_launches=plain))
    batch.render(0, 2)
    device.synchronize()
    for sc, sd, f in zip(scenes, data, frames):
        got = sc.readback(f.width, f.height)
        one = dmt.Scene(device, sd)
        one.render_frames(f, 2)
        want = one.readback(f.width, f.height)
        one.close()
        assert np.array_equal(got, want)
    batch.close()
    for sc in scenes:
        sc.close()


def _direct(device, sd, f, frames):
    sc = dmt.Scene(device, sd)
    sc.render_frames(f, frames)
    out = sc.readback(f.width, f.height)
    sc.close()
    return out


def test_engine_push_renders_the_gradient_sky_lights(device):
    """PtPush sun_and_mode, sun_extra, sun_extra2, moon_dir_phase, moon_extra,
    the denoiser's Halton jitter, and the moon map uploaded as RGBA16F."""
    sd = _gradient_sky_scene()
    f = _sun_moon(sd.frame.copy())
    f.frame_index, f.reset_accum = 3, 1
    dev = rhi.HipDevice(strict=True)
    eng = EngineSim(dev, sd)
    mh, mw = sd.moon.shape[:2]
    moon = dev.create_texture(mw, mh, RGBA16F, STORAGE)
    assert dev.write_texture(moon, np.ascontiguousarray(sd.moon, np.float32).astype(np.float16))
    ov = {"sun_and_mode": ("4f", (*f.sun[:3], 0.0)), "sun_extra": ("4f", tuple(f.sun_extra)),
          "sun_extra2": ("4f", (f.sun_disc_brightness, 0.0, 0.0, 0.0)),
          "moon_dir_phase": ("4f", tuple(f.moon)), "moon_extra": ("4f", tuple(f.moon_extra)),
          "denoiser_enabled": ("I", 1), "halton_jitter": ("2f", (0.25, -0.125))}
    eng.frames(f, 2, textures={7: moon}, overrides=ov)
    assert dev.last_error() == "", dev.last_error()
    got = eng.accum_image()
    dev.close()
    g = f.copy()
    g.flags |= dmt.DMT_FRAME_FIXED_JITTER
    g.jitter[0], g.jitter[1] = 0.25, -0.125
    assert np.array_equal(got, _direct(device, sd, g, 2))
    # the same frames without the moon texture: a different image
    sd.moon = None
    assert not np.array_equal(got, _direct(device, sd, g, 2))


def test_engine_push_renders_the_hdri_cluster_lights(device):
    """PtPush hdri_lights_count / _dir / _col and env_total_luminance, the
    engine's marginal and conditional CDFs at buffer slots 4 and 5."""
    sd, f = _hdri_scene()
    f.frame_index, f.reset_accum = 1, 1
    m, c, total = sd.env_cdf
    dev = rhi.HipDevice(strict=True)
    eng = EngineSim(dev, sd)
    bm = dev.create_buffer(m.nbytes)
    dev.write_buffer(bm, np.ascontiguousarray(m, np.float32))
    bc = dev.create_buffer(c.nbytes)
    dev.write_buffer(bc, np.ascontiguousarray(c, np.float32))
    ov = {"hdri_lights_count": ("I", f.hdri_light_count), "env_total_luminance": ("f", total)}
    for i in range(f.hdri_light_count):
        ov[f"hdri_dir{i}"] = ("4f", tuple(f.hdri_dir[i]))
        ov[f"hdri_col{i}"] = ("4f", tuple(f.hdri_col[i]))
    eng.frames(f, 2, buffers={4: bm, 5: bc}, overrides=ov)
    assert dev.last_error() == "", dev.last_error()
    got = eng.accum_image()
    dev.close()
    assert np.array_equal(got, _direct(device, sd, f, 2))


def test_adjoint_refuses_sky_frames(device):
    sd = _gradient_sky_scene()
    f = _sun_moon(sd.frame.copy())
    sc = dmt.Scene(device, sd)
    with pytest.raises(dmt.DmtError) as e:
        adjoint_numpy(sc, f, np.ones((f.height, f.width, 4), np.float32), device)
    assert e.value.status == dmt.DMT_ERR_UNSUPPORTED
    # without the sky lights the same scene differentiates
    g = adjoint_numpy(sc, sd.frame.copy(), np.ones((f.height, f.width, 4), np.float32), device)
    assert np.all(np.isfinite(g))
    sc.close()
    assert not device.is_lost()
