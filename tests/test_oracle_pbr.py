"""The oracle's PBR material atlas (PathTrace.slang:1435-1605 applyPbrTextures,
:2104-2158 the analytic UV maps, :2536-2543 the mesh UVs, :5913-5919 and
:6162-6171 the ray cone), pinned by closed forms and by exact equivalences:
texel-centre samples are the texels, midpoints their means, taps wrap at
every level, each mip level reads its own region of the band, the sRGB decode
is IEC 61966-2-1, the normal map's basis is orthonormal around n, the sphere,
plane and mesh UVs and the cone's LOD are the closed forms, and a tile of one
constant colour renders exactly as the flat material it scales."""
import math

import numpy as np
import pytest

import dmt
import oracle_ctypes as orc
from scenes import (MAT_LAMBERT, MAT_METAL, NO_TILE, PBR_ATLAS_W, PBR_TILE, atlas, atlas_band, frame, plane,
                    point_light, quad_light, scene, set_tiles, sphere)

RNG = np.random.default_rng(26)


def _oracle(sd):
    return orc.OracleScene.from_scene_data(sd)


def _noise_tile():
    return RNG.integers(0, 256, size=(PBR_TILE, PBR_TILE, 4), dtype=np.uint8)


def _textured_plane_scene(tiles_kw, at, extra=()):
    p = set_tiles(plane((0, 1, 0), 1.0, albedo=(0.8, 0.7, 0.6)), **tiles_kw)
    sd = scene([p, *extra], [point_light((0, 4, 0), (30, 30, 30))])
    sd.atlas = at
    return sd


def test_texel_centres_midpoints_and_wrap():
    base = _noise_tile()
    sd = _textured_plane_scene({"albedo": 1}, atlas([base]))
    o = _oracle(sd)
    k = np.float32(1.0 / 255.0)
    tex = base.astype(np.float32) * k
    xs = RNG.integers(0, PBR_TILE, 64)
    ys = RNG.integers(0, PBR_TILE, 64)
    # texel centres read the texel exactly
    uvl = np.stack([(xs + 0.5) / PBR_TILE, (ys + 0.5) / PBR_TILE, np.zeros(64)], 1)
    got = o.pbr_sample(1, uvl)
    assert np.array_equal(got, tex[ys, xs])
    # the midpoint between two columns is their mean; at column 255 the +1 tap
    # is column 0 (the level wraps)
    xs2 = np.concatenate([xs[:32], np.full(4, PBR_TILE - 1)])
    ys2 = ys[:36]
    uvl = np.stack([(xs2 + 1.0) / PBR_TILE, (ys2 + 0.5) / PBR_TILE, np.zeros(36)], 1)
    got = o.pbr_sample(1, uvl)
    want = 0.5 * (tex[ys2, xs2].astype(np.float64) + tex[ys2, (xs2 + 1) % PBR_TILE])
    assert np.max(np.abs(got - want)) <= 1e-6
    # whole periods of u and v, and negative coordinates, read the same texels
    uvl = np.stack([(xs + 0.5) / PBR_TILE, (ys + 0.5) / PBR_TILE, np.zeros(64)], 1)
    for du, dv in ((1.0, 0.0), (-3.0, 2.0), (0.0, -1.0)):
        shifted = uvl + np.array([du, dv, 0.0])
        assert np.array_equal(o.pbr_sample(1, shifted), tex[ys, xs])


def test_each_mip_level_reads_its_own_region():
    # tile 2's band: level k filled with 10 + 25 k (bypassing the box filter),
    # so a wrong origin or size of a level's region reads another constant
    band = np.zeros((PBR_TILE, PBR_ATLAS_W, 4), np.uint8)
    band[:, :PBR_TILE] = 10
    for lvl in range(1, 9):
        dim = PBR_TILE >> lvl
        y0 = PBR_TILE - 2 * dim
        band[y0:y0 + dim, PBR_TILE:PBR_TILE + dim] = 10 + 25 * lvl
    at = np.concatenate([atlas_band(np.full((PBR_TILE, PBR_TILE, 4), 77, np.uint8)), atlas([])[:PBR_TILE], band])
    sd = _textured_plane_scene({"albedo": 2}, at)
    o = _oracle(sd)
    uv = RNG.random((16, 2))
    k = np.float32(1.0 / 255.0)
    for lvl in range(0, 9):
        got = o.pbr_sample(2, np.column_stack([uv, np.full(16, lvl)]))
        assert np.all(got == np.float32(10 + 25 * lvl) * k), lvl
    # fractional LODs blend the two levels; LODs clamp to [0, 8]
    for lod in (0.25, 3.5, 7.75):
        l0 = int(math.floor(lod))
        a, b = (10 + 25 * l0) / 255.0, (10 + 25 * (l0 + 1)) / 255.0
        got = o.pbr_sample(2, np.column_stack([uv, np.full(16, lod)]))
        assert np.max(np.abs(got - (a + (lod - l0) * (b - a)))) <= 1e-6
    assert np.all(o.pbr_sample(2, np.column_stack([uv, np.full(16, 20.0)])) == np.float32(210) * k)
    assert np.all(o.pbr_sample(2, np.column_stack([uv, np.full(16, -4.0)])) == np.float32(10) * k)


def test_srgb_decode_is_the_iec_eotf():
    c = np.linspace(0.0, 1.0, 4097, dtype=np.float32)
    got = orc.pbr_srgb(c).astype(np.float64)
    cd = c.astype(np.float64)
    want = np.where(cd <= 0.04045, cd / 12.92, ((cd + 0.055) / 1.055) ** 2.4)
    assert np.max(np.abs(got - want) / np.maximum(want, 1e-3)) <= 2e-6
    assert orc.pbr_srgb(np.array([0.0, 1.0], np.float32))[1] == pytest.approx(1.0, abs=2e-7)


def test_normal_map_basis_is_orthonormal_around_n():
    n = RNG.normal(size=(256, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    n[0] = (0, 0, 1)
    n[1] = (0, 0, -1)
    n[2] = (0, 0, -0.99999994)   # near the branch of the basis
    n = n.astype(np.float32)
    ex = orc.pbr_normal_map(n, np.tile([1, 0, 0], (256, 1)))
    ey = orc.pbr_normal_map(n, np.tile([0, 1, 0], (256, 1)))
    ez = orc.pbr_normal_map(n, np.tile([0, 0, 1], (256, 1)))
    assert np.max(np.abs(ez - n)) <= 2e-7
    for a, b in ((ex, ey), (ex, n), (ey, n)):
        assert np.max(np.abs(np.sum(a * b, axis=1))) <= 2e-6
    assert np.max(np.abs(np.cross(ex, ey) - n)) <= 2e-6     # right-handed
    # a general tangent-space vector lands at its components in that basis
    ts = RNG.normal(size=(256, 3)).astype(np.float32)
    ts[:, 2] = np.abs(ts[:, 2]) + 0.1
    got = orc.pbr_normal_map(n, ts)
    want = ex * ts[:, :1] + ey * ts[:, 1:2] + n * ts[:, 2:]
    want /= np.linalg.norm(want, axis=1, keepdims=True)
    assert np.max(np.abs(got - want)) <= 4e-6


def test_sphere_uv_density_and_plane_uvs():
    c, R = np.array([0.5, -0.25, 0.75]), 2.0
    sp = set_tiles(sphere(c, R), albedo=1)
    sd = scene([sp], [point_light((0, 9, 0), (1, 1, 1))])
    sd.atlas = atlas([(128, 128, 128, 255)])
    o = _oracle(sd)
    f = frame((0, 0, 10), (0, 0, 0), 40, 8, 8)
    lat = np.radians([0, 0, 0, 30, -45, 60, 10])
    lon = np.radians([0, 90, 180, -60, 135, 20, -170])
    sn = np.stack([np.cos(lat) * np.cos(lon), np.sin(lat), np.cos(lat) * np.sin(lon)], 1)
    p = c + R * sn
    orig = p + 5.0 * sn
    s = o.pbr_surface(f, orig, -sn)
    assert np.all(s[:, 13] == 1)
    u = 0.5 + np.arctan2(sn[:, 2], sn[:, 0]) / (2 * math.pi)
    v = 0.5 - np.arcsin(sn[:, 1]) / math.pi
    assert np.max(np.abs(s[:, 0] - u)) <= 3e-6
    assert np.max(np.abs(s[:, 1] - v)) <= 3e-6
    cos_lat = np.maximum(np.cos(lat), 0.05)
    scale = np.maximum(1 / (math.pi * R), 1 / (2 * math.pi * R * cos_lat))
    assert np.max(np.abs(s[:, 2] / scale - 1)) <= 2e-6
    # planes: y-facing -> (x, z), x-facing -> (z, y), z-facing -> (x, y), the
    # dominant axis of a tilted normal deciding; one UV per metre
    for n_pl, want in (((0, 1, 0), (0, 2)), ((1, 0, 0), (2, 1)), ((0, 0, 1), (0, 1)), ((0.6, 0.8, 0), (0, 2)),
                       ((0.8, 0.0, 0.6), (2, 1))):
        npl = np.array(n_pl, np.float64)
        pl = set_tiles(plane(npl, 1.5), albedo=1)
        sdp = scene([pl], [])
        sdp.atlas = sd.atlas
        op = _oracle(sdp)
        pts = RNG.uniform(-3, 3, size=(8, 3))
        pts -= (pts @ npl + 1.5)[:, None] * npl      # onto n.p + d = 0
        s = op.pbr_surface(f, pts + 4.0 * npl, np.tile(-npl, (8, 1)))
        assert np.all(s[:, 13] == 1)
        assert np.max(np.abs(s[:, 0] - pts[:, want[0]])) <= 2e-5
        assert np.max(np.abs(s[:, 1] - pts[:, want[1]])) <= 2e-5
        assert np.all(s[:, 2] == 1.0)


def test_cone_lod_closed_form():
    sd = _textured_plane_scene({"albedo": 1}, atlas([_noise_tile()]))
    o = _oracle(sd)
    f = frame((0, 0, 10), (0, 0, 0), 40, 8, 8)
    ang = np.radians([0, 30, 60, 80, 89.9])
    d = np.stack([np.sin(ang), -np.cos(ang), np.zeros_like(ang)], 1)
    origin = np.tile([0.0, 2.0, 0.0], (ang.size, 1))
    for w in (0.0, 1e-4, 3e-3, 0.05, 2.0):
        s = o.pbr_surface(f, origin, d, cone_width=w)
        cos_i = np.maximum(np.abs(np.cos(ang)), 1e-3)
        texels = np.maximum(w / cos_i * 1.0 * PBR_TILE, 1.0)
        want = np.clip(np.log2(texels), 0, 8)
        assert np.max(np.abs(s[:, 3] - want)) <= 2e-5, w


def test_mesh_uv_is_the_barycentric_interpolation():
    v = np.array([[-2, -1, -2], [3, -1, -1], [-1, -1, 3]], np.float32)
    uvs = np.array([[0.1, 0.2], [1.3, 0.2], [0.1, -0.9]], np.float32)
    sd = scene([], [point_light((0, 4, 0), (10, 10, 10))], positions=v, indices=[[0, 1, 2]])
    sd.atlas = atlas([_noise_tile()])
    sd.uvs = uvs
    o = _oracle(sd)
    f = frame((0, 3, 0), (0, -1, 0.01), 40, 8, 8)
    f.flags |= dmt.FRAME_MESH_TEXTURES
    f.mesh_tex[:] = (1, NO_TILE, NO_TILE, NO_TILE)
    ab = RNG.uniform(0.05, 0.45, size=(16, 2))
    pts = v[0] + ab[:, :1] * (v[1] - v[0]) + ab[:, 1:] * (v[2] - v[0])
    s = o.pbr_surface(f, pts + np.array([0.1, 2.0, -0.2]), np.tile([-0.1, -2.0, 0.2], (16, 1)) / np.linalg.norm([0.1, 2, 0.2]))
    assert np.all(s[:, 13] == 2)
    want = uvs[0] + ab[:, :1] * (uvs[1] - uvs[0]) + ab[:, 1:] * (uvs[2] - uvs[0])
    assert np.max(np.abs(s[:, 0:2] - want)) <= 2e-5
    assert np.all(s[:, 2] == 0.0) and np.all(s[:, 3] == 0.0)    # meshes read level 0
    # the gold metal's albedo scaled by the decoded texture
    assert not np.allclose(s[:, 4:7], [1.0, 0.85, 0.45])
    # without the frame flag the mesh stays untextured
    f.flags &= ~dmt.FRAME_MESH_TEXTURES
    s = o.pbr_surface(f, pts + np.array([0.1, 2.0, -0.2]), np.tile([-0.1, -2.0, 0.2], (16, 1)) / np.linalg.norm([0.1, 2, 0.2]))
    assert np.allclose(s[:, 4:7], [1.0, 0.85, 0.45], atol=0)


def _c0_like(seed=3):
    sd = dmt.generate(0, seed=seed, width=48, height=32)
    sd.frame.spp = 2
    return sd


def _render(sd):
    img, cnt = _oracle(sd).render(sd.frame, 32, 0)
    return img, cnt


@pytest.mark.parametrize("byte", [0, 37, 128, 255])
def test_constant_albedo_tile_renders_as_the_scaled_flat_albedo(byte):
    sd = _c0_like()
    flat = _c0_like()
    x = orc.pbr_srgb(np.array([np.float32(byte) * np.float32(1.0 / 255.0)], np.float32))[0]
    # glass uses its albedo only in shadow-ray transmittance, which reads the
    # flat value (the maps apply at shading vertices, PathTrace.slang:6171),
    # so the flat scene scales the Lambert and metal albedos
    for i in range(sd.prims.shape[0]):
        set_tiles(sd.prims[i], albedo=1)
        if sd.prims[i, 9] != 2:
            flat.prims[i, 4:7] = (sd.prims[i, 4:7].astype(np.float32) * np.float32(x)).astype(np.float32)
    sd.atlas = atlas([(byte, byte, byte, 255)])
    a, ca = _render(sd)
    b, cb = _render(flat)
    assert np.array_equal(a, b) and ca == cb


def test_constant_roughness_and_metallic_tiles_render_as_the_flat_materials():
    sd, flat = _c0_like(5), _c0_like(5)
    r = np.float32(90) * np.float32(1.0 / 255.0)
    for i in range(sd.prims.shape[0]):
        if sd.prims[i, 9] == MAT_METAL:
            set_tiles(sd.prims[i], rough=1)
            flat.prims[i, 7] = np.float32(min(max(np.float32(sd.prims[i, 7] * r), 0.0), 1.0))
        elif sd.prims[i, 9] == MAT_LAMBERT:
            set_tiles(sd.prims[i], metal=2)
            flat.prims[i, 9] = MAT_METAL
    sd.atlas = atlas([(90, 0, 0, 255), (200, 0, 0, 255)])
    a, ca = _render(sd)
    b, cb = _render(flat)
    assert np.array_equal(a, b) and ca == cb
    # a metallic value at or below 0.5 keeps the Lambert material
    sd2, plain = _c0_like(5), _c0_like(5)
    for i in range(sd2.prims.shape[0]):
        set_tiles(sd2.prims[i], metal=1)
    sd2.atlas = atlas([(127, 0, 0, 255)])
    assert np.array_equal(_render(sd2)[0], _render(plain)[0])


def test_untextured_and_out_of_range_tiles_leave_the_image_unchanged():
    plain = _c0_like(7)
    ref = _render(plain)[0]
    sd = _c0_like(7)
    sd.atlas = atlas([_noise_tile()])
    assert np.array_equal(_render(sd)[0], ref)           # atlas, every tile none
    for i in range(sd.prims.shape[0]):
        set_tiles(sd.prims[i], albedo=2, normal=7, rough=2, metal=99)   # the atlas holds tiles 0 and 1
    assert np.array_equal(_render(sd)[0], ref)


def test_textured_render_follows_the_fp64_oracle():
    # at equal spp the fp32 and fp64 renders follow the same random stream:
    # they agree except where a decision flips at a rounding-level boundary
    rng = np.random.default_rng(31)
    sd = _c0_like(9)
    sd.atlas = atlas([rng.integers(0, 256, size=(PBR_TILE, PBR_TILE, 4), dtype=np.uint8) for _ in range(2)])
    for i in range(sd.prims.shape[0]):
        set_tiles(sd.prims[i], albedo=1, normal=2)
    img = _render(sd)[0][..., :3].astype(np.float64)
    assert not np.array_equal(img, _render(_c0_like(9))[0][..., :3])
    f64 = _oracle(sd).render_f64(sd.frame)
    d = np.abs(img - f64).max(-1) / np.maximum(np.abs(f64).max(-1), 1e-3)
    flipped = d > 1e-3
    assert flipped.mean() <= 0.02
    keep = ~flipped
    assert np.sum((img[keep] - f64[keep]) ** 2) / np.sum(f64[keep] ** 2) <= 1e-7
