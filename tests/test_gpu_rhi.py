"""The RHI adapter on the GPU: the engine's call sequence through
dmt::rhi::HipDevice renders the same accumulation buffer, bit for bit, as the C
ABI called directly (itself pinned to the CPU oracle by the parity tests), for
scenes of every kind; the output texture holds the display transform of the
mean; material edits take the no-rebuild path, geometry edits rebuild and keep
the running mean; TLAS instances flatten with their transforms; unsupported
features are refused (strict) or left out (lenient) without touching the
image.
"""
import numpy as np
import pytest
import torch

import _dmt_rhi as rhi
import dmt
import oracle_ctypes as orc
from rhi_engine import RGBA32F, STORAGE, EngineSim, pack_ptpush, tonemap_reference

pytestmark = pytest.mark.gpu


def _scenes():
    out = []
    sd = dmt.generate(0, seed=11, width=72, height=48)
    out.append(("spheres+plane, zero env", sd))
    sd = dmt.generate(0, seed=12, width=64, height=40)
    sd.env = None
    out.append(("spheres+plane, gradient sky", sd))
    sd = dmt.generate(1, seed=13, subdivisions=4, random_tris=3000, width=80, height=45)
    out.append(("mesh + lognormal env", sd))
    sd = dmt.generate(2, seed=14, spheres=200, subdivisions=3, random_tris=4000, lights=128, width=64, height=36)
    out.append(("many lights, light tree", sd))
    sd = dmt.generate(4, seed=15, spheres=40, width=48, height=48)
    sd.frame.spp = 4
    out.append(("spp 4", sd))
    return out


SCENES = _scenes()


def _direct(device, sd, frame, count, materials=None):
    sc = dmt.Scene(device, sd)
    if materials is not None:
        sc.set_materials(materials)
    sc.render_frames(frame, count)
    img = sc.readback(frame.width, frame.height)
    sc.close()
    return img


@pytest.mark.parametrize("device_bvh", [False, True], ids=["host-bvh", "device-bvh"])
@pytest.mark.parametrize("idx", range(len(SCENES)), ids=[s[0] for s in SCENES])
def test_engine_frames_equal_direct_render(device, idx, device_bvh):
    label, sd = SCENES[idx]
    f = sd.frame.copy()
    f.frame_index = 3
    f.reset_accum = 1
    dev = rhi.HipDevice(device_bvh_build=device_bvh)
    assert not dev.is_device_lost(), dev.last_error()
    eng = EngineSim(dev, sd)
    eng.frames(f, 3)
    got = eng.accum_image()
    assert dev.last_error() == "", dev.last_error()
    assert dev.scene_builds() == 1 and dev.material_updates() == 0
    want = _direct(device, sd, f, 3)
    same = float(np.mean(np.all(got == want, axis=-1)))
    print(f"{label}: engine sequence == direct render on {same:.6f} of pixels")
    assert np.array_equal(got, want), label
    assert np.all(got[..., 3] == 3.0)
    dev.close()


def test_engine_frame_matches_the_oracle(device):
    sd = dmt.generate(0, seed=21, width=40, height=30)
    f = sd.frame.copy()
    f.reset_accum = 1
    dev = rhi.HipDevice()
    eng = EngineSim(dev, sd)
    eng.frame(f)
    ref, _ = orc.OracleScene.from_scene_data(sd).render(f, 32, 0)
    assert np.array_equal(eng.accum_image(), ref)
    dev.close()


@pytest.mark.parametrize("op", [0, 1, 2, 3, 4, 9])
def test_output_texture_is_the_display_transform(device, op):
    label, sd = SCENES[2]
    f = sd.frame.copy()
    f.reset_accum = 1
    exposure = 0.0 if op == 9 else 1.7   # exposure <= 0 reads as 1
    dev = rhi.HipDevice()
    eng = EngineSim(dev, sd, exposure=exposure)
    out = eng.frames(f, 2, tonemap_op=op)
    acc = eng.accum_image()
    img = dev.readback_texture(out)
    want = tonemap_reference(acc[..., :3], exposure if exposure > 0 else 1.0, op if op <= 4 else 0)
    err = float(np.max(np.abs(img[..., :3].astype(np.float64) - want)))
    print(f"op {op}: max |output - display transform of the mean| = {err:.2e}")
    assert err <= 2e-5
    assert np.all(img[..., 3] == 1.0)
    dev.close()


def test_material_edit_takes_the_no_rebuild_path(device):
    label, sd = SCENES[0]
    f = sd.frame.copy()
    f.reset_accum = 1
    dev = rhi.HipDevice()
    eng = EngineSim(dev, sd)
    eng.frames(f, 2)
    prims = sd.prims.copy()
    rng = np.random.default_rng(5)
    prims[:, 4:7] = rng.uniform(0.1, 0.9, (prims.shape[0], 3)).astype(np.float32)   # albedo
    prims[:, 7] = rng.uniform(0.0, 0.5, prims.shape[0]).astype(np.float32)          # roughness
    prims[3, 16:19] = np.float32(4.0)                                                # emission
    dev.write_buffer(eng.prims, prims)
    g = f.copy()
    g.frame_index = f.frame_index + 2
    eng.frame(g)
    assert dev.scene_builds() == 1 and dev.material_updates() == 1, dev.last_error()
    edited = dmt.SceneData(prims=prims, lights=sd.lights, positions=sd.positions, indices=sd.indices, env=sd.env,
                           env_intensity=sd.env_intensity, frame=sd.frame)
    assert np.array_equal(eng.accum_image(), _direct(device, edited, g, 1))
    dev.close()


def test_geometry_edit_rebuilds_and_keeps_the_running_mean(device):
    label, sd = SCENES[0]
    f = sd.frame.copy()
    f.reset_accum = 1
    w, h = f.width, f.height
    dev = rhi.HipDevice()
    eng = EngineSim(dev, sd)
    eng.frames(f, 2)
    moved = sd.prims.copy()
    sphere = int(np.nonzero(moved[:, 8] == 0.0)[0][0])
    moved[sphere, 0:3] += np.float32(0.75)      # v0: a new centre (v3 follows it)
    moved[sphere, 12:15] = moved[sphere, 0:3]
    dev.write_buffer(eng.prims, moved)
    g = f.copy()
    g.frame_index = f.frame_index + 2
    g.reset_accum = 0
    eng.frame(g)
    assert dev.scene_builds() == 2 and dev.material_updates() == 0, dev.last_error()
    got = eng.accum_image()
    # direct: two frames of the old scene, then the moved scene accumulating
    # into the same memory (dmt_scene_bind_accum)
    before = _direct(device, sd, f, 2)
    mem = torch.from_numpy(before.reshape(-1).copy()).to("cuda:0")
    sd2 = dmt.SceneData(prims=moved, lights=sd.lights, positions=sd.positions, indices=sd.indices, env=sd.env,
                        env_intensity=sd.env_intensity, frame=sd.frame)
    sc = dmt.Scene(device, sd2)
    sc.bind_accum(mem.data_ptr(), w, h)
    sc.render(g)
    want = sc.readback(w, h)
    assert np.array_equal(mem.cpu().numpy().reshape(h, w, 4), want)   # the bound memory is the image
    sc.bind_accum(0, 0, 0)
    sc.close()
    assert np.array_equal(got, want)
    assert np.all(got[..., 3] == 3.0)
    dev.close()


def test_tlas_instances_flatten_with_their_transforms(device):
    sd = dmt.generate(1, seed=31, subdivisions=3, random_tris=400, width=64, height=40)
    pos, idx = sd.positions.astype(np.float32), sd.indices.astype(np.uint32)
    m = np.array([0.8, -0.3, 0.1, 1.5, 0.25, 0.9, -0.2, -0.4, -0.1, 0.3, 0.95, 0.7], np.float32)
    dev = rhi.HipDevice()
    eng = EngineSim(dev, dmt.SceneData(prims=sd.prims, lights=sd.lights, env=sd.env, env_intensity=sd.env_intensity,
                                       frame=sd.frame))
    blas = dev.create_blas(pos, idx)
    eng.tlas = dev.create_tlas([(blas, None, 0, 0xFF), (blas, list(m), 1, 0xFF), (blas, list(m), 2, 0)])
    f = sd.frame.copy()
    f.reset_accum = 1
    eng.frames(f, 2)
    got = eng.accum_image()
    assert dev.last_error() == "", dev.last_error()
    # the flattened mesh in float32, the same operation order as the adapter
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    t = np.stack([((m[0] * x + m[1] * y) + m[2] * z) + m[3], ((m[4] * x + m[5] * y) + m[6] * z) + m[7],
                  ((m[8] * x + m[9] * y) + m[10] * z) + m[11]], axis=1).astype(np.float32)
    flat = dmt.SceneData(prims=sd.prims, lights=sd.lights, positions=np.concatenate([pos, t]),
                         indices=np.concatenate([idx, idx + np.uint32(pos.shape[0])]), env=sd.env,
                         env_intensity=sd.env_intensity, frame=sd.frame)
    assert np.array_equal(got, _direct(device, flat, f, 2))
    dev.close()


def test_strict_refusal_and_lenient_rendering(device):
    label, sd = SCENES[0]
    f = sd.frame.copy()
    f.reset_accum = 1
    restir = {"restir_enabled": ("I", 1)}
    strict = rhi.HipDevice(strict=True)
    eng = EngineSim(strict, sd)
    eng.frame(f)
    before = eng.accum_image()
    g = f.copy()
    g.frame_index += 1
    g.reset_accum = 0
    eng.frame(g, overrides=restir)
    assert "ReSTIR" in strict.last_error()
    assert np.array_equal(eng.accum_image(), before)          # a refused dispatch is a no-op
    eng.frame(g, push=eng.push(g)[:100])                       # short push: no-op too
    assert "shorter than the 112 B" in strict.last_error()
    assert np.array_equal(eng.accum_image(), before)
    assert not strict.is_device_lost()
    strict.close()
    lenient = rhi.HipDevice(strict=False)
    eng = EngineSim(lenient, sd)
    eng.frames(f, 2, overrides=restir)
    assert "ReSTIR" in lenient.ignored_features()
    assert np.array_equal(eng.accum_image(), _direct(device, sd, f, 2))
    lenient.close()


def test_dispatch_errors_leave_the_device_usable(device):
    label, sd = SCENES[0]
    f = sd.frame.copy()
    f.reset_accum = 1
    dev = rhi.HipDevice()
    eng = EngineSim(dev, sd)
    # more primitives than the buffer holds
    eng.frame(f, overrides={"prim_count": ("I", sd.prims.shape[0] + 1)})
    assert "buffer slot 3" in dev.last_error()
    # the swapchain resized but not the accumulation texture
    dev.resize(f.width + 8, f.height)
    eng.frame(f)
    assert "differ" in dev.last_error()
    dev.resize(f.width, f.height)
    # no accumulation texture bound
    dev.bind_texture(1, 0)
    dev.push_constants(eng.push(f))
    dev.dispatch()
    assert "slot 1" in dev.last_error()
    # other pipelines are no-ops
    clear = dev.create_pipeline("clear")
    dev.bind_pipeline(clear)
    dev.dispatch()
    assert dev.scene_builds() == 0
    eng.frames(f, 2)
    assert np.array_equal(eng.accum_image(), _direct(device, sd, f, 2))
    # clear / copy into a texture, readback full-or-nothing
    t = dev.create_texture(4, 2, RGBA32F, STORAGE)
    dev.clear_texture(t, [0.25, 0.5, 0.75, 1.0])
    assert np.array_equal(dev.readback_texture(t), np.tile(np.float32([0.25, 0.5, 0.75, 1.0]), (2, 4, 1)))
    src = dev.create_buffer(8 * 16)
    dev.write_buffer(src, np.arange(32, dtype=np.float32))
    dev.copy_buffer_to_texture(src, t)
    assert np.array_equal(dev.readback_texture(t).reshape(-1), np.arange(32, dtype=np.float32))
    dev.destroy_texture(eng.accum)                             # unbinds it from the scene first
    assert not dev.is_device_lost()
    dev.wait_idle()
    dev.close()


def test_bind_accum_rejects_bad_memory(device):
    label, sd = SCENES[0]
    f = sd.frame.copy()
    sc = dmt.Scene(device, sd)
    host = np.zeros(f.width * f.height * 4, np.float32)
    with pytest.raises(dmt.DmtError):
        sc.bind_accum(host.ctypes.data, f.width, f.height)      # host memory
    mem = torch.zeros(f.width * f.height * 4 + 1, dtype=torch.float32, device="cuda:0")
    with pytest.raises(dmt.DmtError):
        sc.bind_accum(mem.data_ptr() + 4, f.width, f.height)    # misaligned
    sc.bind_accum(mem.data_ptr(), f.width, f.height)
    g = f.copy()
    g.width += 1
    with pytest.raises(dmt.DmtError):
        sc.render(g)                                            # another resolution
    sc.render(f)
    torch.cuda.synchronize()
    assert float(mem[3].item()) == 1.0                          # one sample in the bound memory
    sc.bind_accum(0, 0, 0)
    sc.render(g)                                                # its own buffer again, any size
    assert sc.readback(g.width, g.height)[..., 3].min() == 1.0
    sc.close()
