"""Water through the C ABI, against the oracle bit for bit: a wavy water
surface over the sphere workload seen from above (submerged spheres and
floor, shadow rays through the surface) and from below (the camera in the
water, the deep-water tint), water with SDF clusters under the sun, water
with PBR maps, depth of field, motion blur and dispersion; the engine's push
(water_params0/1, water_state) equals the C ABI's render; a batch with water
scenes equals the scenes rendered alone; the adjoint refuses water, invalid
water parameters are refused."""
import numpy as np
import pytest
import torch

import dmt
import oracle_ctypes as orc
from scenes import (MAT_LAMBERT, MAT_METAL, NO_TILE, PBR_TILE, SDF_BOX, SDF_LEAF, SDF_SPHERE, atlas,
                    batch_against_scenes, plane, sdf_cluster, sdf_node, set_tiles)

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(134)


def _run(device, sd, f, frames=2):
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


def _check(label, got, ref, counts, oc):
    same = float(np.mean(np.all(got == ref, axis=-1)))
    print(f"{label}: equal to the oracle on {same:.6f} of pixels", flush=True)
    assert np.array_equal(got, ref), label
    assert (counts["primary_rays"], counts["extension_rays"], counts["shadow_rays"]) == (
        oc["primary"], oc["extension"], oc["shadow"]), label


def _water(f, amp=0.35):
    f.water_absorb[:] = (0.08, 0.15, 0.3)
    f.water_ior = 1.33
    f.wave[:] = (1.2, amp, 0.6, 3.0)
    f.deep_color[:] = (0.01, 0.03, 0.06)
    return f


def _lake(seed=21, w=56, h=40, spp=2, level=-1.0):
    sd = dmt.generate(0, seed=seed, width=w, height=h)
    # the surface at y = level (its own index 0: the frame's)
    sd.prims = np.vstack([sd.prims, plane((0, 1, 0), -level, albedo=(1, 1, 1), mat=dmt.MAT_WATER, ior=0.0)])
    f = _water(sd.frame.copy())
    f.spp = spp
    return sd, f


def test_wavy_surface_seen_from_above(device):
    sd, f = _lake()
    f = dmt.look_at((0.0, 3.0, 11.0), (0.0, -2.0, 0.0), 55.0, f.width, f.height, base=f)
    got, ref, counts, oc = _run(device, sd, f)
    _check("water from above", got, ref, counts, oc)
    assert counts["shadow_rays"] > 0


def test_camera_in_the_water(device):
    sd, f = _lake(seed=22)
    f = dmt.look_at((0.5, -2.5, 7.0), (0.0, -1.0, 0.0), 60.0, f.width, f.height, base=f)
    f.camera_in_water = 1
    got, ref, counts, oc = _run(device, sd, f)
    _check("camera in the water", got, ref, counts, oc)


def test_water_with_sdf_clusters_under_the_sun(device):
    sd, f = _lake(seed=23, w=48, h=36)
    sd.env = None
    sd.sdf = np.stack([
        sdf_cluster([sdf_node(SDF_LEAF, SDF_BOX, params=(0.5, 0.7, 0.4, 0), center=(-2.0, -1.0, 2.0))],
                    (-2.55, -1.75, 1.55), (-1.45, -0.25, 2.45), mat=MAT_LAMBERT, albedo=(0.7, 0.6, 0.5)),
        sdf_cluster([sdf_node(SDF_LEAF, SDF_SPHERE, params=(0.6, 0, 0, 0), center=(2.0, -1.2, 1.5))],
                    (1.35, -1.85, 0.85), (2.65, -0.55, 2.15), mat=MAT_METAL, albedo=(0.9, 0.9, 0.9), rough=0.1),
    ])
    f = dmt.look_at((0.0, 2.0, 10.0), (0.0, -1.5, 0.0), 55.0, 48, 36, base=f)
    fwd = np.array(f.fwd_aspect[:3], np.float64)
    sun = fwd + np.array([0.2, 1.2, 0.1])
    sun /= np.linalg.norm(sun)
    f.sun[:] = (*sun, 0.0)
    f.sun_extra[:] = (2.0, 1.5, 0.7, 0.0)
    f.sun_disc_brightness = 0.5
    f.sdf_max_iters, f.sdf_epsilon = 128, 1e-4
    got, ref, counts, oc = _run(device, sd, f)
    _check("water + SDF + sun", got, ref, counts, oc)


def test_sdf_cluster_of_water(device):
    """A rounded box of material 3 (an SDF pool) over a lit floor: its
    surface bends by the waves and refracts into absorbing water."""
    sd = dmt.generate(0, seed=27, width=40, height=30)
    sd.sdf = np.stack([
        sdf_cluster([sdf_node(SDF_LEAF, SDF_BOX, params=(2.5, 0.6, 2.0, 0.1), center=(0.0, -0.6, 1.0))],
                    (-2.65, -1.35, -1.15), (2.65, 0.15, 3.15), mat=3, albedo=(1, 1, 1), ior=1.33),
    ])
    f = _water(sd.frame.copy(), amp=0.25)
    f = dmt.look_at((0.0, 3.0, 10.0), (0.0, -1.0, 1.0), 50.0, 40, 30, base=f)
    f.spp = 2
    f.sdf_max_iters, f.sdf_epsilon = 128, 1e-4
    got, ref, counts, oc = _run(device, sd, f)
    _check("SDF cluster of water", got, ref, counts, oc)


def test_water_with_maps_lens_motion_blur_and_dispersion(device):
    sd, f = _lake(seed=24, w=48, h=36, spp=3)
    sd.atlas = atlas([RNG.integers(0, 256, size=(PBR_TILE, PBR_TILE, 4), dtype=np.uint8)])
    for i in range(sd.prims.shape[0] - 1):
        set_tiles(sd.prims[i], albedo=1)
        if sd.prims[i, 8] == 0:
            sd.prims[i, 12:15] = sd.prims[i, 0:3] + RNG.uniform(-0.2, 0.2, 3)
    f = dmt.look_at((0.0, 3.0, 11.0), (0.0, -2.0, 0.0), 55.0, 48, 36, base=f)
    f.dof[:] = (0.06, 11.0, 5.0, 0.2)
    f.motion_blur = 1
    f.spectral = 1
    f.cauchy_c = 0.0042
    got, ref, counts, oc = _run(device, sd, f)
    _check("water + maps + DOF + motion blur + spectral", got, ref, counts, oc)


def test_engine_push_renders_water(device):
    """water_params0 (912), water_params1 (928), water_state (1040) as
    Engine.cpp:8883-8890 and :9263-9266 fill them; the water plane is a
    primitive of material 3 at buffer slot 3."""
    import _dmt_rhi as rhi
    from rhi_engine import EngineSim

    sd, f = _lake(seed=25, w=40, h=30)
    f = dmt.look_at((0.0, 3.0, 11.0), (0.0, -2.0, 0.0), 55.0, 40, 30, base=f)
    sd.frame = f.copy()
    sc = dmt.Scene(device, sd)
    sc.render_frames(f, 2)
    want = sc.readback(f.width, f.height)
    sc.close()
    dev = rhi.HipDevice(strict=True)
    eng = EngineSim(dev, sd)
    ov = {"water_params0": ("4f", (*f.water_absorb, f.water_ior)), "water_params1": ("4f", tuple(f.wave)),
          "water_state": ("4f", (0.0, *f.deep_color))}
    eng.frames(f, 2, overrides=ov)
    assert dev.last_error() == "", dev.last_error()
    assert np.array_equal(eng.accum_image(), want)
    # the FFT ocean is refused (strict) as a no-op
    eng.frames(f, 1, overrides={**ov, "ocean_params0": ("4f", (1.0, 256.0, 0.0, 0.0))})
    assert "FFT ocean" in dev.last_error()
    dev.close()


def test_water_in_a_batch_equals_the_scenes_rendered_alone(device):
    """A batch of a lake seen from above, a camera in the water, an SDF
    cluster of water with PBR maps on the spheres, and a dry scene runs the
    batched kernel's water variants; each accumulation equals the scene
    rendered alone bit for bit, as do the ray counts."""
    sd1, f1 = _lake(seed=28, w=32, h=24)
    f1 = dmt.look_at((0.0, 3.0, 11.0), (0.0, -2.0, 0.0), 55.0, 32, 24, base=f1)
    sd2, f2 = _lake(seed=29, w=24, h=24, spp=3)
    f2 = dmt.look_at((0.5, -2.5, 7.0), (0.0, -1.0, 0.0), 60.0, 24, 24, base=f2)
    f2.camera_in_water = 1
    sd3 = dmt.generate(0, seed=30, width=32, height=16)
    sd3.sdf = np.stack([
        sdf_cluster([sdf_node(SDF_LEAF, SDF_BOX, params=(2.5, 0.6, 2.0, 0.1), center=(0.0, -0.6, 1.0))],
                    (-2.65, -1.35, -1.15), (2.65, 0.15, 3.15), mat=3, albedo=(1, 1, 1), ior=1.33),
    ])
    sd3.atlas = atlas([RNG.integers(0, 256, size=(PBR_TILE, PBR_TILE, 4), dtype=np.uint8)])
    for i in range(sd3.prims.shape[0]):
        set_tiles(sd3.prims[i], albedo=1)
    f3 = _water(sd3.frame.copy(), amp=0.25)
    f3.spp = 2
    f3.sdf_max_iters, f3.sdf_epsilon = 128, 1e-4
    sd4 = dmt.generate(0, seed=31, width=16, height=16)
    cases = [(sd1, f1), (sd2, f2), (sd3, f3), (sd4, sd4.frame)]
    alone, batched, c_alone, c_batch = batch_against_scenes(device, cases)
    for k, (a, b) in enumerate(zip(alone, batched)):
        assert np.array_equal(a, b), k
    assert c_alone == c_batch


def test_water_is_refused_by_the_adjoint_and_checked(device):
    sd, f = _lake(seed=26, w=32, h=24)
    sc = dmt.Scene(device, sd)
    grads = torch.zeros(sd.prims.shape[0] * dmt.MATERIAL_FLOATS, dtype=torch.float32, device="cuda")
    dimg = torch.ones(f.width * f.height * 4, dtype=torch.float32, device="cuda")
    with pytest.raises(dmt.DmtError) as e:
        sc.adjoint(f, dimg.data_ptr(), grads.data_ptr())
    assert e.value.status == dmt.DMT_ERR_UNSUPPORTED
    for field, value in (("water_absorb", (-0.1, 0.0, 0.0)), ("deep_color", (0.0, -1.0, 0.0)),
                         ("wave", (np.nan, 0.0, 0.0, 0.0)), ("camera_in_water", 2)):
        g = f.copy()
        if isinstance(value, tuple):
            getattr(g, field)[:] = value
        else:
            setattr(g, field, value)
        with pytest.raises(dmt.DmtError) as e:
            sc.render(g)
        assert e.value.status == dmt.DMT_ERR_INVALID_ARG, field
    sc.close()
