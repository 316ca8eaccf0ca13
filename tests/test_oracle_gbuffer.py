"""The oracle's G-buffer pass (the denoiser inputs of PathTrace.slang:5463-5648)
pinned by what the camera and the geometry fix: the current projection puts
every primary point at its pixel centre, a still camera has zero motion, a
translated camera moves points by the float64 reprojection, the ground plane's
depth is its closed-form distance along the forward axis, sphere normals point
from the centre, and the specular fields follow the material table."""
import numpy as np
import pytest

import dmt
import oracle_ctypes as orc
from rhi_engine import pixel_dirs, to_pixel, view_proj


def _plane_scene(mat, rough, width=40, height=30):
    prim = np.zeros((1, 28), np.float32)
    prim[0, 0:4] = (0.0, 1.0, 0.0, 5.0)           # n . p + 5 = 0: the plane y = -5
    prim[0, 4:8] = (0.3, 0.5, 0.7, rough)
    prim[0, 8:12] = (1.0, float(mat), 1.5, 0.0)
    prim[0, 12:16] = prim[0, 0:4]
    prim[0, 20:24] = (0.0, 0.0, 0.0, 1.0)
    frame = dmt.look_at((0.0, 2.0, 12.0), (0.0, -1.0, 0.0), 50.0, width, height)
    frame.spp, frame.max_bounces = 1, 2
    return dmt.SceneData(prims=prim, frame=frame)


@pytest.fixture(scope="module")
def c0():
    sd = dmt.generate(0, seed=31, width=48, height=36)
    return sd, orc.OracleScene.from_scene_data(sd)


def test_still_camera_has_no_motion(c0):
    sd, osc = c0
    M = view_proj(sd.frame).astype(np.float32)
    g = osc.gbuffer(sd.frame, M, M)
    assert np.all(g[..., 1:3] == 0.0)
    assert np.all(g[..., 14:16] == 0.0)


def test_current_projection_puts_points_at_pixel_centres(c0):
    sd, osc = c0
    f = sd.frame
    M = view_proj(f).astype(np.float32)
    fixed = np.zeros(16, np.float32)
    fixed[15] = 1.0          # clip = (0, 0, 0, 1) for every point: the image centre
    g = osc.gbuffer(f, M, fixed)
    cur = np.stack([f.width / 2.0 - g[..., 1], f.height / 2.0 - g[..., 2]], axis=-1)
    x, y = np.meshgrid(np.arange(f.width) + 0.5, np.arange(f.height) + 0.5)
    err = np.abs(cur - np.stack([x, y], axis=-1)).max()
    print(f"max |current pixel - pixel centre| = {err:.2e} px")
    assert err < 2e-3
    # a transposed matrix puts them elsewhere
    gt = osc.gbuffer(f, M.reshape(4, 4).T.reshape(16).copy(), fixed)
    assert np.abs(gt[..., 1:3] - g[..., 1:3]).max() > 1.0


def test_translated_camera_moves_points_by_the_reprojection(c0):
    sd, osc = c0
    f = sd.frame
    Mc = view_proj(f)
    shift = np.array(f.right_xyz[:3]) * 0.3 + np.array(f.up_xyz[:3]) * 0.1
    Mp = view_proj(f, pos=np.array(f.pos_fovtan[:3]) + shift)
    g = osc.gbuffer(f, Mc.astype(np.float32), Mp.astype(np.float32))
    d = pixel_dirs(f)
    fwd = np.array(f.fwd_aspect[:3]) / np.linalg.norm(f.fwd_aspect[:3])
    hit = g[..., 0] < 1e9
    t = np.where(hit, g[..., 0] / (d @ fwd), 1000.0)
    world = np.array(f.pos_fovtan[:3]) + d * t[..., None]
    want = to_pixel(Mp, world, f.width, f.height) - to_pixel(Mc, world, f.width, f.height)
    err = np.abs(g[..., 1:3] - want).max()
    print(f"max |motion - float64 reprojection| = {err:.2e} px; motion up to {np.abs(want).max():.1f} px")
    assert np.abs(want).max() > 1.0 and err < 5e-3


@pytest.mark.parametrize("mat,rough", [(0, 0.3), (1, 0.3), (2, 0.0), (1, 1.7)])
def test_ground_plane_fields(mat, rough):
    sd = _plane_scene(mat, rough)
    f = sd.frame
    osc = orc.OracleScene.from_scene_data(sd)
    M = view_proj(f).astype(np.float32)
    g = osc.gbuffer(f, M, M)
    d = pixel_dirs(f)
    py = float(f.pos_fovtan[1])
    down = d[..., 1] < -1e-4
    t = np.where(down, -(py + 5.0) / np.where(down, d[..., 1], -1.0), np.inf)
    hit = g[..., 0] < 1e9
    assert np.array_equal(hit, np.isfinite(t) & (t > 1e-3))
    fwd = np.array(f.fwd_aspect[:3]) / np.linalg.norm(f.fwd_aspect[:3])
    depth = t * (d @ fwd)
    assert np.allclose(g[..., 0][hit], depth[hit], rtol=2e-5)
    assert np.all(g[..., 0][~hit] == np.float32(1e10))
    assert np.all(g[hit][:, 3:6] == np.float32([0.0, 1.0, 0.0]))
    assert np.all(g[hit][:, 6:9] == np.float32([0.3, 0.5, 0.7]))
    f0 = {0: [0.0, 0.0, 0.0], 1: [0.3, 0.5, 0.7], 2: [0.04, 0.04, 0.04]}[mat]
    assert np.all(g[hit][:, 9:12] == np.float32(f0))
    r = min(max(rough, 0.0), 1.0)
    assert np.all(g[hit][:, 12] == np.float32(1.0 if mat == 0 else r))
    assert np.allclose(g[hit][:, 13], t[hit] * (1.0 - r), rtol=2e-5, atol=1e-6)
    # misses: the sky sentinels
    assert np.all(g[~hit][:, 3:12] == 0.0) and np.all(g[~hit][:, 12] == 1.0) and np.all(g[~hit][:, 13] == 0.0)


def test_sphere_normals_point_from_the_centre():
    prim = np.zeros((1, 28), np.float32)
    prim[0, 0:4] = (0.0, 0.0, 0.0, 1.0)
    prim[0, 4:8] = (0.8, 0.8, 0.8, 0.0)
    prim[0, 8:12] = (0.0, 0.0, 1.0, 0.0)
    prim[0, 12:16] = prim[0, 0:4]
    prim[0, 20:24] = (0.0, 0.0, 0.0, 1.0)
    f = dmt.look_at((0.0, 0.0, 5.0), (0.0, 0.0, 0.0), 40.0, 33, 33)
    sd = dmt.SceneData(prims=prim, frame=f)
    osc = orc.OracleScene.from_scene_data(sd)
    M = view_proj(f).astype(np.float32)
    g = osc.gbuffer(f, M, M)
    hit = g[..., 0] < 1e9
    assert hit[16, 16] and not hit[0, 0]
    d = pixel_dirs(f)
    fwd = np.array(f.fwd_aspect[:3])
    world = np.array(f.pos_fovtan[:3]) + d * (g[..., 0] / (d @ fwd))[..., None]
    n = world / np.linalg.norm(world, axis=-1, keepdims=True)
    assert np.abs(g[hit][:, 3:6] - n[hit]).max() < 1e-5
    assert np.abs(g[16, 16, 3:6] - [0.0, 0.0, 1.0]).max() < 1e-6
    assert np.abs(g[16, 16, 0] - 4.0) < 1e-5          # depth of the nearest point


def test_mesh_hits_carry_the_gold_metal():
    sd = dmt.generate(1, seed=32, subdivisions=3, random_tris=300, width=40, height=30)
    osc = orc.OracleScene.from_scene_data(sd)
    M = view_proj(sd.frame).astype(np.float32)
    g = osc.gbuffer(sd.frame, M, M)
    mesh = np.all(g[..., 9:12] == np.float32([1.0, 0.85, 0.45]), axis=-1)   # PathTrace.slang:2652-2658
    assert mesh.sum() > 100
    assert np.all(g[mesh][:, 6:9] == np.float32([1.0, 0.85, 0.45]))
    assert np.all(g[mesh][:, 12] == np.float32(0.05))
    d = pixel_dirs(sd.frame)
    t = g[..., 0] / (d @ np.array(sd.frame.fwd_aspect[:3]))
    assert np.allclose(g[mesh][:, 13], t[mesh] * 0.95, rtol=1e-5)
