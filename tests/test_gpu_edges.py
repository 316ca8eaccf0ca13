"""Edge cases through the C ABI, each against the oracle: an empty scene (sky
only, with and without the env map), a scene of lights alone, one-pixel and
one-row/one-column frames, a single sphere filling the view, and the size
limits."""
import numpy as np
import pytest

import dmt
import oracle_ctypes as orc

pytestmark = pytest.mark.gpu


def _frame(w, h, spp=1, bounces=4):
    f = dmt.look_at((0.0, 2.0, 12.0), (0.0, 0.0, 0.0), 50.0, w, h)
    f.spp, f.max_bounces, f.reset_accum = spp, bounces, 1
    f.mis_enabled, f.light_tree_enabled, f.refract_bounces, f.firefly_clamp = 1, 1, 4, 10.0
    return f


def _render_both(device, sd, f, frames=2):
    sc = dmt.Scene(device, sd)
    sc.render_frames(f, frames)
    got = sc.readback(f.width, f.height)
    sc.close()
    osc = orc.OracleScene.from_scene_data(sd)
    acc = None
    for k in range(frames):
        g = f.copy()
        g.frame_index = f.frame_index + k
        g.reset_accum = 1 if k == 0 else 0
        acc, _ = osc.render(g, 32, 0, accum=acc)
    return got, acc


def _env(seed=3, w=32, h=16):
    return np.exp(np.random.default_rng(seed).normal(0.0, 1.0, (h, w, 3))).astype(np.float32)


@pytest.mark.parametrize("env", [False, True], ids=["gradient sky", "env map"])
def test_empty_scene_is_the_sky(device, env):
    sd = dmt.SceneData(env=_env() if env else None, frame=_frame(24, 16))
    got, ref = _render_both(device, sd, sd.frame)
    assert np.array_equal(got, ref)
    assert np.all(got[..., :3] > 0.0)


def test_lights_without_geometry(device):
    base = dmt.generate(2, seed=5, spheres=10, subdivisions=1, random_tris=10, lights=32, width=8, height=8)
    sd = dmt.SceneData(lights=base.lights, env=_env(), frame=_frame(20, 12))
    got, ref = _render_both(device, sd, sd.frame)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("w,h", [(1, 1), (1, 37), (53, 1), (9, 7), (65, 3)])
def test_tiny_and_ragged_frames(device, w, h):
    sd = dmt.generate(0, seed=8, width=8, height=8)
    f = sd.frame.copy()
    f.width, f.height = w, h
    f.fwd_aspect = (f.fwd_aspect[0], f.fwd_aspect[1], f.fwd_aspect[2], float(w) / float(h))
    f.spp = 3
    got, ref = _render_both(device, sd, f)
    assert np.array_equal(got, ref)


def test_sphere_filling_the_view(device):
    prim = np.zeros((1, 28), np.float32)
    prim[0, 0:4] = (0.0, 0.0, 0.0, 50.0)      # the camera sits inside a large sphere
    prim[0, 4:8] = (0.6, 0.6, 0.6, 0.0)
    prim[0, 8:12] = (0.0, 0.0, 1.5, 0.0)
    prim[0, 12:16] = prim[0, 0:4]
    prim[0, 20:24] = (0.0, 0.0, 0.0, 1.0)
    base = dmt.generate(0, seed=1, width=8, height=8)
    sd = dmt.SceneData(prims=prim, lights=base.lights, frame=_frame(32, 20))
    got, ref = _render_both(device, sd, sd.frame)
    assert np.array_equal(got, ref)


def test_size_limits(device):
    sd = dmt.generate(0, seed=1, width=8, height=8)
    sc = dmt.Scene(device, sd)
    for w, h in ((65536, 1), (1, 65536), (8193, 8193)):    # > 65535 per side, > 2^26 pixels
        f = sd.frame.copy()
        f.width, f.height = w, h
        with pytest.raises(dmt.DmtError) as e:
            sc.render(f)
        assert e.value.status == dmt.DMT_ERR_LIMIT, (w, h)
    f = sd.frame.copy()
    f.spp = 33
    with pytest.raises(dmt.DmtError):
        sc.render(f)
    f = sd.frame.copy()
    f.refract_bounces = 65
    with pytest.raises(dmt.DmtError) as e:
        sc.render(f)
    assert e.value.status == dmt.DMT_ERR_LIMIT
    sc.close()
    assert not device.is_lost()
