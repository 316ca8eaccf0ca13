"""Camera-sample features through the C ABI, against the oracle bit for bit:
thin lens (round, polygonal, cat's-eye), motion blur (moving spheres, a
translated mesh), hero-wavelength dispersion and all of them together; the
G-buffer of a translated mesh; the engine path (PtPush fields decoded by the
RHI device) equal to the direct render; a batch mixing them equal to the
per-scene renders; and invalid feature values refused.  (The adjoint with
the features: tests/test_gpu_adjoint.py.)"""
import numpy as np
import pytest
import torch

import _dmt_rhi as rhi
import dmt
import oracle_ctypes as orc
from rhi_engine import EngineSim, view_proj

pytestmark = pytest.mark.gpu


def _moving(sd, seed=11, amount=0.6):
    """The scene with every sphere given a previous centre (v3) off its current one."""
    prims = sd.prims.copy()
    rng = np.random.default_rng(seed)
    spheres = prims[:, 8] == 0.0
    prims[spheres, 12:15] = prims[spheres, 0:3] - rng.uniform(-amount, amount, (int(spheres.sum()), 3))
    out = dmt.SceneData(prims=prims, lights=sd.lights, positions=sd.positions, indices=sd.indices, env=sd.env,
                        env_intensity=sd.env_intensity, frame=sd.frame.copy())
    return out


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


def _check(label, got, ref):
    same = float(np.mean(np.all(got == ref, axis=-1)))
    print(f"{label}: equal to the oracle on {same:.6f} of pixels")
    assert np.array_equal(got, ref), label


def _spheres():
    return dmt.generate(0, seed=17, width=48, height=36)


def _mesh():
    """A small mesh workload (icosphere + random triangles, env map) with the
    spheres, plane and quad light of the sphere workload."""
    m = dmt.generate(1, seed=42, subdivisions=3, random_tris=500, width=40, height=30)
    s = dmt.generate(0, seed=19, width=8, height=8)
    return dmt.SceneData(prims=s.prims, lights=s.lights, positions=m.positions, indices=m.indices, env=m.env,
                         env_intensity=m.env_intensity, frame=m.frame.copy())


LENSES = [("round", (0.25, 11.0, 0.0, 0.0), 0.0), ("hexagon", (0.3, 9.5, 6.0, 0.3), 0.0),
          ("pentagon + cat's eye", (0.35, 12.0, 5.0, 1.1), 0.7), ("round + cat's eye", (0.2, 10.0, 1.0, 0.0), 1.0)]


@pytest.mark.parametrize("idx", range(len(LENSES)), ids=[l[0] for l in LENSES])
def test_thin_lens_equals_the_oracle(device, idx):
    label, dof, cat = LENSES[idx]
    sd = _spheres()
    f = sd.frame.copy()
    f.spp = 2
    f.dof = dof
    f.cat_eye = cat
    got, ref = _render_both(device, sd, f)
    _check(label, got, ref)
    plain, _ = _render_both(device, sd, sd.frame)
    assert not np.array_equal(got, plain)


@pytest.mark.parametrize("which", ["spheres", "mesh"])
def test_motion_blur_equals_the_oracle(device, which):
    sd = _moving(_spheres() if which == "spheres" else _mesh())
    f = sd.frame.copy()
    f.spp = 2
    f.motion_blur = 1
    if which == "mesh":
        f.mesh_shift_active = 1
        f.mesh_shift_prev = (-0.4, 0.1, 0.0)
        f.mesh_shift_curr = (0.3, 0.0, 0.2)
    got, ref = _render_both(device, sd, f)
    _check(which, got, ref)


def test_translated_mesh_without_motion_blur(device):
    """mesh_motion_prev.w on, r_motion_blur off: the mesh sits at its current
    translation (the shutter time is 1)."""
    sd = _mesh()
    f = sd.frame.copy()
    f.mesh_shift_active = 1
    f.mesh_shift_prev = (5.0, 5.0, 5.0)   # ignored at time 1
    f.mesh_shift_curr = (0.5, -0.25, 0.0)
    got, ref = _render_both(device, sd, f)
    _check("translated mesh", got, ref)


def test_motion_blur_off_ignores_previous_centres(device):
    sd = _moving(_spheres())
    still = _spheres()
    got, _ = _render_both(device, sd, sd.frame)
    want, _ = _render_both(device, still, still.frame)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("cauchy", [0.0042, 0.03])
def test_spectral_dispersion_equals_the_oracle(device, cauchy):
    sd = _spheres()
    f = sd.frame.copy()
    f.spp = 3
    f.spectral = 1
    f.cauchy_c = cauchy
    got, ref = _render_both(device, sd, f)
    _check(f"spectral C={cauchy}", got, ref)


def test_spectral_one_spp_keeps_signed_frame_values(device):
    """At 1 spp a spectral frame value carries the hero wavelength's signed
    sRGB weight; negative channels reach the running mean unclamped
    (PathTrace.slang:7722-7734), in the frame-completion kernel and in the
    batched kernel alike."""
    sd = _spheres()
    f = sd.frame.copy()
    f.spp = 1
    f.spectral = 1
    f.cauchy_c = 0.0042
    got1, ref1 = _render_both(device, sd, f, frames=1)
    _check("spectral 1 spp, one frame", got1, ref1)
    assert (got1[..., :3] < 0.0).any()
    got, ref = _render_both(device, sd, f, frames=4)
    _check("spectral 1 spp, four frames", got, ref)
    sc = dmt.Scene(device, sd)
    batch = dmt.Batch(device, [sc], [f])
    batch.render(0, 4)
    device.synchronize()
    assert np.array_equal(sc.readback(f.width, f.height), ref)
    batch.close()
    sc.close()


def test_all_features_together(device):
    sd = _moving(_mesh(), seed=23)
    f = sd.frame.copy()
    f.spp = 3
    f.dof = (0.2, 10.0, 7.0, 0.2)
    f.cat_eye = 0.5
    f.motion_blur = 1
    f.spectral = 1
    f.cauchy_c = 0.01
    f.mesh_shift_active = 1
    f.mesh_shift_prev = (0.0, 0.0, 0.0)
    f.mesh_shift_curr = (0.25, 0.0, -0.25)
    got, ref = _render_both(device, sd, f, frames=3)
    _check("dof + motion + spectral + mesh shift", got, ref)
    st = dmt.Scene(device, sd)
    counts = st.measure_traffic(f)
    st.close()
    _, oc = orc.OracleScene.from_scene_data(sd).render(f, 32, 0)
    assert (counts["primary_rays"], counts["extension_rays"], counts["shadow_rays"]) == (
        oc["primary"], oc["extension"], oc["shadow"])


def test_gbuffer_of_a_translated_mesh(device):
    sd = _mesh()
    f = sd.frame.copy()
    f.mesh_shift_active = 1
    f.mesh_shift_prev = (9.0, 9.0, 9.0)
    f.mesh_shift_curr = (0.4, 0.0, -0.3)
    f.motion_blur = 1
    vc = view_proj(f).astype(np.float32)
    n = f.width * f.height
    bufs = {k: torch.full((n * 4,), -7.0, dtype=torch.float32, device="cuda:0") for k in ("depth", "normal")}
    torch.cuda.synchronize()
    sc = dmt.Scene(device, sd)
    sc.gbuffer(f, vc, vc, 4, **{k: v.data_ptr() for k, v in bufs.items()})
    device.synchronize()
    sc.close()
    want = orc.OracleScene.from_scene_data(sd).gbuffer(f, vc, vc)
    depth = bufs["depth"].cpu().numpy().reshape(f.height, f.width, 4)[..., 0]
    normal = bufs["normal"].cpu().numpy().reshape(f.height, f.width, 4)[..., :3]
    assert np.array_equal(depth, want[..., 0])
    assert np.array_equal(normal, want[..., 3:6])
    g = f.copy()
    g.mesh_shift_active = 0
    assert not np.array_equal(orc.OracleScene.from_scene_data(sd).gbuffer(g, vc, vc)[..., 0], want[..., 0])


def test_engine_push_renders_the_features(device):
    sd = _moving(_spheres(), seed=29)
    f = sd.frame.copy()
    f.frame_index = 2
    f.reset_accum = 1
    f.dof = (0.2, 11.0, 6.0, 0.1)
    f.cat_eye = 0.4
    f.motion_blur = 1
    f.spectral = 1
    f.cauchy_c = 0.0042
    dev = rhi.HipDevice(strict=True)
    eng = EngineSim(dev, sd)
    eng.frames(f, 2)
    assert dev.last_error() == "", dev.last_error()
    got = eng.accum_image()
    dev.close()
    sc = dmt.Scene(device, sd)
    sc.render_frames(f, 2)
    want = sc.readback(f.width, f.height)
    sc.close()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("plain", [0, 1], ids=["graph", "plain launches"])
def test_batch_with_features_equals_per_scene_renders(device, plain):
    """A batch whose scenes use different features (none; lens + motion blur;
    dispersion + a translated mesh) equals each scene rendered on its own."""
    a = _spheres()
    b = _moving(_spheres(), seed=31)
    c = _mesh()
    fa = a.frame.copy()
    fb = b.frame.copy()
    fb.spp, fb.dof, fb.cat_eye, fb.motion_blur = 2, (0.3, 9.5, 6.0, 0.3), 0.5, 1
    fc = c.frame.copy()
    fc.spectral, fc.cauchy_c, fc.mesh_shift_active = 1, 0.02, 1
    fc.mesh_shift_prev, fc.mesh_shift_curr = (0.0, 0.0, 0.0), (0.3, 0.1, 0.0)
    scenes = [dmt.Scene(device, sd) for sd in (a, b, c)]
    frames = [fa, fb, fc]
    batch = dmt.Batch(device, scenes, frames, dmt.BatchParams(plain_launches=plain))
    batch.render(0, 2)
    device.synchronize()
    for sc, sd, f in zip(scenes, (a, b, c), frames):
        got = sc.readback(f.width, f.height)
        one = dmt.Scene(device, sd)
        one.render_frames(f, 2)
        want = one.readback(f.width, f.height)
        one.close()
        assert np.array_equal(got, want)
    _, ref = _render_both(device, b, fb)
    assert np.array_equal(scenes[1].readback(fb.width, fb.height), ref)
    batch.close()
    for sc in scenes:
        sc.close()


def test_invalid_feature_values_are_refused(device):
    sd = _spheres()
    sc = dmt.Scene(device, sd)
    cases = [("dof", (float("nan"), 1.0, 0.0, 0.0), dmt.DMT_ERR_INVALID_ARG),
             ("dof", (0.1, 5.0, 100.0, 0.0), dmt.DMT_ERR_LIMIT),
             ("motion_blur", 2, dmt.DMT_ERR_INVALID_ARG),
             ("mesh_shift_curr", (0.0, float("inf"), 0.0), dmt.DMT_ERR_INVALID_ARG),
             ("reserved", 1, dmt.DMT_ERR_INVALID_ARG)]
    for field, value, status in cases:
        bad = sd.frame.copy()
        setattr(bad, field, value)
        with pytest.raises(dmt.DmtError) as e:
            sc.render(bad)
        assert e.value.status == status, field
    sc.close()
    assert not device.is_lost()
