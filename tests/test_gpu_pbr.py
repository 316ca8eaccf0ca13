"""PBR material atlas through the C ABI, against the oracle bit for bit:
albedo, normal, roughness and metallic maps on spheres and planes (noise
tiles, so every level and both bilinear taps matter, at grazing angles where
the ray cone selects coarse levels), several samples per pixel (the cone
restarts per sample), the UV-mapped mesh with the frame's mesh tiles, maps
together with SDF clusters and the sun, and with depth of field, motion blur
and dispersion; removing the atlas restores the untextured image; a batch of
textured and untextured scenes equals the scenes rendered alone; the adjoint
refuses textured frames, and malformed atlases are refused."""
import numpy as np
import pytest
import torch

import dmt
import oracle_ctypes as orc
from scenes import (MAT_DIELECTRIC, MAT_LAMBERT, MAT_METAL, NO_TILE, SDF_BOX, SDF_LEAF, SDF_SPHERE, PBR_TILE,
                    atlas, batch_against_scenes, sdf_cluster, sdf_node, set_tiles)

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(2026)


def _noise_tiles(n):
    return [RNG.integers(0, 256, size=(PBR_TILE, PBR_TILE, 4), dtype=np.uint8) for _ in range(n)]


def _smooth_normal_tile():
    # a gentle bump map: tangent-space normals near (0, 0, 1)
    y, x = np.mgrid[0:PBR_TILE, 0:PBR_TILE] / PBR_TILE
    nx = 0.35 * np.sin(2 * np.pi * 3 * x)
    ny = 0.35 * np.cos(2 * np.pi * 2 * y)
    nz = np.sqrt(np.maximum(1 - nx * nx - ny * ny, 0.0))
    t = np.stack([nx, ny, nz, np.ones_like(nx)], -1) * 0.5 + 0.5
    return np.clip(np.round(t * 255), 0, 255).astype(np.uint8)


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


def _textured_spheres(seed=11, w=64, h=48, spp=2):
    sd = dmt.generate(0, seed=seed, width=w, height=h)
    tiles = _noise_tiles(3) + [_smooth_normal_tile()]
    sd.atlas = atlas(tiles)
    # a mix: albedo maps everywhere, roughness on metals, metallic on a third
    # of the Lambert spheres, normal maps (noise and smooth) on half of them;
    # the ground plane (prim 0) gets the albedo map and the smooth normal map
    for i in range(sd.prims.shape[0]):
        mat = int(sd.prims[i, 9])
        kw = {"albedo": 1 + (i % 2)}
        if mat == MAT_METAL:
            kw["rough"] = 3
        if mat == MAT_LAMBERT and i % 3 == 0 and i > 0:
            kw["metal"] = 3
        if i % 2 == 0:
            kw["normal"] = 4 if i % 4 == 0 else 3
        set_tiles(sd.prims[i], **kw)
    f = dmt.look_at((0.0, 1.0, 12.0), (0.0, -3.0, 0.0), 55.0, w, h, base=sd.frame)   # the ground plane recedes
    f.spp = spp
    return sd, f


def test_textured_spheres_and_plane_equal_the_oracle(device):
    sd, f = _textured_spheres()
    got, ref, counts, oc = _run(device, sd, f)
    _check("textured spheres + plane", got, ref, counts, oc)
    plain = dmt.generate(0, seed=11, width=64, height=48)
    sc = dmt.Scene(device, plain)
    sc.render_frames(f, 2)
    assert not np.array_equal(sc.readback(f.width, f.height), got)
    sc.close()


def test_many_samples_per_pixel_restart_the_cone(device):
    sd, f = _textured_spheres(seed=12, w=40, h=32, spp=5)
    got, ref, counts, oc = _run(device, sd, f, frames=1)
    _check("5 spp textured", got, ref, counts, oc)


def test_uv_mapped_mesh(device):
    m = dmt.generate(1, seed=5, subdivisions=3, random_tris=400, width=48, height=36)
    uvs = RNG.uniform(-1.5, 2.5, size=(m.positions.shape[0], 2)).astype(np.float32)
    sd = dmt.SceneData(prims=m.prims, lights=m.lights, positions=m.positions, indices=m.indices, env=m.env,
                       env_intensity=m.env_intensity, frame=m.frame.copy(), uvs=uvs,
                       atlas=atlas(_noise_tiles(2) + [_smooth_normal_tile()]))
    f = sd.frame.copy()
    f.spp = 2
    f.flags |= dmt.FRAME_MESH_TEXTURES
    f.mesh_tex[:] = (1, 3, 2, NO_TILE)
    got, ref, counts, oc = _run(device, sd, f)
    _check("UV-mapped mesh", got, ref, counts, oc)
    # the frame flag off: the mesh renders untextured (and so does the scene
    # without the atlas)
    f.flags &= ~dmt.FRAME_MESH_TEXTURES
    got2, ref2, _, _ = _run(device, sd, f, frames=1)
    assert np.array_equal(got2, ref2)
    sc = dmt.Scene(device, sd)
    sc.set_pbr_atlas(None)
    sc.render(f)
    assert np.array_equal(sc.readback(f.width, f.height), got2)
    sc.close()


def test_maps_with_sdf_clusters_and_the_sun(device):
    sd, f = _textured_spheres(seed=13, w=48, h=36)
    sd.env = None
    clusters = [
        sdf_cluster([sdf_node(SDF_LEAF, SDF_BOX, params=(0.5, 0.4, 0.3, 0), center=(-2.5, 0.5, 2.0))],
                    (-3.05, -0.05, 1.65), (-1.95, 1.05, 2.35), mat=MAT_LAMBERT, albedo=(0.7, 0.6, 0.5)),
        sdf_cluster([sdf_node(SDF_LEAF, SDF_SPHERE, params=(0.6, 0, 0, 0), center=(2.5, 0.3, 2.0))],
                    (1.85, -0.35, 1.35), (3.15, 0.95, 2.65), mat=MAT_METAL, albedo=(0.9, 0.9, 0.9), rough=0.1),
    ]
    sd.sdf = np.stack(clusters)
    fwd = np.array(f.fwd_aspect[:3], np.float64)
    sun = fwd + np.array([0.2, 0.9, 0.1])
    sun /= np.linalg.norm(sun)
    f.sun[:] = (*sun, 0.0)
    f.sun_extra[:] = (2.0, 1.5, 0.7, 0.0)
    f.sun_disc_brightness = 0.5
    f.sdf_max_iters, f.sdf_epsilon = 128, 1e-4
    got, ref, counts, oc = _run(device, sd, f)
    _check("maps + SDF clusters + sun", got, ref, counts, oc)


def test_maps_with_lens_motion_blur_and_dispersion(device):
    sd, f = _textured_spheres(seed=14, w=48, h=36, spp=3)
    for i in range(sd.prims.shape[0]):
        if sd.prims[i, 8] == 0:
            sd.prims[i, 12:15] = sd.prims[i, 0:3] + RNG.uniform(-0.2, 0.2, 3)   # previous centres
    f.dof[:] = (0.08, 11.0, 6.0, 0.3)
    f.cat_eye = 0.25
    f.motion_blur = 1
    f.spectral = 1
    f.cauchy_c = 0.0042
    got, ref, counts, oc = _run(device, sd, f)
    _check("maps + DOF + motion blur + spectral", got, ref, counts, oc)


def test_textured_frames_in_a_batch_equal_the_scenes_rendered_alone(device):
    """A batch of two textured scenes (different atlases, sizes and sample
    counts) and an untextured one runs the batched kernel's PBR variant; each
    scene's accumulation equals its render alone bit for bit, as do the ray
    counts."""
    sd1, f1 = _textured_spheres(seed=15, w=32, h=24)
    sd2, f2 = _textured_spheres(seed=16, w=40, h=24, spp=3)
    sd3 = dmt.generate(0, seed=17, width=24, height=16)
    cases = [(sd1, f1), (sd2, f2), (sd3, sd3.frame)]
    alone, batched, c_alone, c_batch = batch_against_scenes(device, cases)
    for k, (a, b) in enumerate(zip(alone, batched)):
        assert np.array_equal(a, b), k
    assert c_alone == c_batch


def test_textured_frames_are_refused_by_the_adjoint(device):
    sd, f = _textured_spheres(seed=15, w=32, h=24)
    sc = dmt.Scene(device, sd)
    grads = torch.zeros(sd.prims.shape[0] * dmt.MATERIAL_FLOATS, dtype=torch.float32, device="cuda")
    dimg = torch.ones(f.width * f.height * 4, dtype=torch.float32, device="cuda")
    with pytest.raises(dmt.DmtError) as e:
        sc.adjoint(f, dimg.data_ptr(), grads.data_ptr())
    assert e.value.status == dmt.DMT_ERR_UNSUPPORTED
    # malformed atlases
    for shape in ((256, 383, 4), (100, 384, 4), (0, 384, 4)):
        with pytest.raises(dmt.DmtError) as e:
            sc.set_pbr_atlas(np.zeros(shape, np.uint8))
        assert e.value.status == dmt.DMT_ERR_INVALID_ARG
    sc.close()


def test_engine_sequence_with_the_atlas_and_mesh_uvs(device):
    """The engine's binds: the atlas as an RGBA8 texture at storage slot 16
    (Engine.cpp:3984-4020, :7216), the mesh UVs at buffer slot 17 (:7137-7150),
    the prims' tiles in v6 and mesh_tex_indices in the push (:9131-9134); the
    accumulation equals the C ABI's render of the same scene and frames."""
    import _dmt_rhi as rhi
    from rhi_engine import NO_TEX_TILE, RGBA8_UNORM, SAMPLED, EngineSim

    base, f = _textured_spheres(seed=16, w=40, h=30)
    mesh = dmt.generate(1, seed=6, subdivisions=2, random_tris=120)
    pos = (mesh.positions * 0.6 + np.array([0.0, -1.0, 1.5], np.float32)).astype(np.float32)
    uvs = RNG.uniform(-1.0, 2.0, size=(pos.shape[0], 2)).astype(np.float32)
    sd = dmt.SceneData(prims=base.prims, lights=base.lights, positions=pos, indices=mesh.indices, env=base.env,
                       env_intensity=base.env_intensity, frame=f.copy(), atlas=base.atlas, uvs=uvs)
    f.flags |= dmt.FRAME_MESH_TEXTURES
    f.mesh_tex[:] = (2, 4, 3, NO_TEX_TILE)
    sc = dmt.Scene(device, sd)
    sc.render_frames(f, 2)
    want = sc.readback(f.width, f.height)
    sc.close()

    dev = rhi.HipDevice(strict=True)
    eng = EngineSim(dev, sd)
    at = dev.create_texture(base.atlas.shape[1], base.atlas.shape[0], RGBA8_UNORM, SAMPLED)
    assert dev.write_texture(at, np.ascontiguousarray(base.atlas))
    ub = dev.create_buffer(uvs.nbytes)
    dev.write_buffer(ub, uvs)
    ov = {"mesh_tex_indices": ("4I", tuple(int(t) for t in f.mesh_tex))}
    eng.frames(f, 2, textures={16: at}, buffers={17: ub}, overrides=ov)
    assert dev.last_error() == "", dev.last_error()
    assert np.array_equal(eng.accum_image(), want)
    # mesh tiles with a UV buffer short of the mesh's vertices are refused;
    # without the atlas the frame renders untextured, as the C ABI renders the
    # scene without one
    short = dev.create_buffer(16)
    dev.write_buffer(short, np.zeros(4, np.float32))
    eng.frames(f, 1, textures={16: at}, buffers={17: short}, overrides=ov)
    assert "buffer slot 17" in dev.last_error()
    plain = dmt.Scene(device, dmt.SceneData(prims=base.prims, lights=base.lights, positions=pos,
                                            indices=mesh.indices, env=base.env, env_intensity=base.env_intensity,
                                            frame=f.copy()))
    g = f.copy()
    g.flags &= ~dmt.FRAME_MESH_TEXTURES
    plain.render_frames(g, 2)
    want_plain = plain.readback(f.width, f.height)
    plain.close()
    dev2 = rhi.HipDevice(strict=True)
    eng2 = EngineSim(dev2, sd)
    eng2.frames(f, 2)
    assert dev2.last_error() == "", dev2.last_error()
    assert np.array_equal(eng2.accum_image(), want_plain)
    dev.close()
    dev2.close()
