"""Helpers that build small scenes in the reference wire formats for tests."""
import math

import numpy as np

import dmt

MAT_LAMBERT, MAT_METAL, MAT_DIELECTRIC = 0, 1, 2


def prim(kind, v0, radius_or_d, albedo=(0.5, 0.5, 0.5), rough=0.0, mat=MAT_LAMBERT, ior=1.0, user_id=0,
         emission=(0, 0, 0)):
    p = np.zeros(28, np.float32)
    p[0:3] = v0
    p[3] = radius_or_d
    p[4:7] = albedo
    p[7] = rough
    p[8] = kind
    p[9] = mat
    p[10] = ior
    p[11] = np.array([user_id], np.uint32).view(np.float32)[0]
    p[12:15] = v0
    p[16:19] = emission
    p[23] = 1.0
    p[24:28] = np.array([0xFFFFFFFF] * 4, np.uint32).view(np.float32)
    return p


def sphere(c, r, **kw):
    return prim(0, c, r, **kw)


def plane(n, d, **kw):
    return prim(1, n, d, **kw)


def light(type_, pos, intensity, radius=0.0, dir=(0, -1, 0), cos_outer=0.0, u=(1, 0, 0), v_half=0.0,
          cos_inner=None):
    L = np.zeros(20, np.float32)
    L[0:3] = pos
    L[3] = radius
    L[4:7] = intensity
    L[7] = type_
    L[8:11] = dir
    L[11] = cos_outer
    L[12:15] = u
    if cos_inner is not None:
        L[12] = cos_inner
    L[15] = v_half
    L[19] = 1.0
    return L


def point_light(pos, I):
    return light(0, pos, I)


def quad_light(pos, n, u, v_half, I):
    return light(3, pos, I, dir=n, u=u, v_half=v_half)


def sphere_light(pos, r, I):
    return light(2, pos, I, radius=r)


def triangle_light(a, b, c, Le):
    """Emissive triangle (light type 4): one-sided along cross(b - a, c - a)."""
    L = np.zeros(20, np.float32)
    L[0:3] = a
    L[4:7] = Le
    L[7] = 4.0
    L[8:11] = b
    L[12:15] = c
    L[19] = 1.0
    return L


def frame(eye, target, fov, w, h, spp=1, bounces=4, frame_index=0):
    f = dmt.look_at(eye, target, fov, w, h)
    f.spp = spp
    f.max_bounces = bounces
    f.frame_index = frame_index
    f.reset_accum = 1
    f.mis_enabled = 1
    f.light_tree_enabled = 1
    f.refract_bounces = 4
    f.firefly_clamp = 10.0
    f.ema_alpha = 0.0
    return f


def scene(prims=(), lights=(), positions=None, indices=None, env=None, env_intensity=1.0, sdf=None):
    return dmt.SceneData(
        prims=np.array(prims, np.float32).reshape(-1, 28),
        lights=np.array(lights, np.float32).reshape(-1, 20),
        positions=np.zeros((0, 3), np.float32) if positions is None else np.asarray(positions, np.float32),
        indices=np.zeros((0, 3), np.uint32) if indices is None else np.asarray(indices, np.uint32),
        env=env, env_intensity=env_intensity,
        sdf=None if sdf is None or len(sdf) == 0 else np.asarray(sdf, np.float32).reshape(-1, 112))


def black_env(w=8, h=4):
    return np.zeros((h, w, 3), np.float32)


def polygon_irradiance(x, n, verts):
    """Irradiance at x (normal n) from a unit-radiance Lambertian polygon
    (Lambert's polygon formula): E = 1/2 sum_i angle_i * n . unit(r_i x r_{i+1})."""
    r = [np.asarray(v, np.float64) - np.asarray(x, np.float64) for v in verts]
    r = [v / np.linalg.norm(v) for v in r]
    e = 0.0
    for i in range(len(r)):
        a, b = r[i], r[(i + 1) % len(r)]
        c = np.cross(a, b)
        e += math.acos(max(-1.0, min(1.0, float(np.dot(a, b))))) * float(np.dot(n, c / np.linalg.norm(c)))
    return abs(0.5 * e)


SDF_SPHERE, SDF_BOX, SDF_ROUND_BOX, SDF_TORUS, SDF_CAPSULE, SDF_PLANE = range(6)
SDF_LEAF, SDF_UNION, SDF_SUBTRACT, SDF_INTERSECT, SDF_DISPLACE = range(5)


def sdf_node(op=SDF_LEAF, shape=0, a=0, b=0, params=(0, 0, 0, 0), center=(0, 0, 0)):
    return (op, shape, a, b, tuple(params), tuple(center))


def sdf_cluster(nodes, lo, hi, mat=MAT_LAMBERT, albedo=(0.5, 0.5, 0.5), rough=0.0, ior=1.0):
    """One SDF cluster record (112 floats, the engine's sdf_clusters layout):
    nodes in post-order as sdf_node tuples, box [lo, hi]."""
    r = np.zeros(112, np.float32)
    u = r.view(np.uint32)
    r[0:3], r[4:7] = lo, hi
    u[3], u[7] = len(nodes), mat
    r[8:11], r[11], r[12] = albedo, rough, ior
    for i, (op, shape, a, b, q, c) in enumerate(nodes):
        o = 16 + 12 * i
        u[o:o + 4] = (op, shape, a, b)
        r[o + 4:o + 8] = q
        r[o + 8:o + 11] = c
    return r


def sdf_box(center, half, pad=1e-3, **kw):
    c, h = np.asarray(center, np.float32), np.asarray(half, np.float32)
    return sdf_cluster([sdf_node(SDF_LEAF, SDF_BOX, params=(*h, 0.0), center=c)], c - h - pad, c + h + pad, **kw)


def sdf_ball(center, r, pad=1e-3, **kw):
    c = np.asarray(center, np.float32)
    return sdf_cluster([sdf_node(SDF_LEAF, SDF_SPHERE, params=(r, 0, 0, 0), center=c)], c - r - pad, c + r + pad, **kw)


NO_TILE = 0xFFFFFFFF
PBR_TILE, PBR_ATLAS_W = 256, 384


def set_tiles(p, albedo=NO_TILE, normal=NO_TILE, rough=NO_TILE, metal=NO_TILE):
    """A primitive's v6: its albedo, normal, roughness and metallic atlas tiles."""
    p[24:28] = np.array([albedo, normal, rough, metal], np.uint32).view(np.float32)
    return p


def atlas_band(base):
    """One 256-row tile band of the engine's atlas (384 x 256 RGBA8): the base
    image on the left, mips 1..8 (2x2 means rounded in byte space) stacked in
    the right column at rows 256 - 2 * dim (Engine::BuildPbrTileMips)."""
    base = np.asarray(base, np.uint8).reshape(PBR_TILE, PBR_TILE, 4)
    band = np.zeros((PBR_TILE, PBR_ATLAS_W, 4), np.uint8)
    band[:, :PBR_TILE] = base
    level = base.astype(np.uint32)
    dim = PBR_TILE
    while dim > 1:
        nd = dim // 2
        s = level[0::2, 0::2] + level[0::2, 1::2] + level[1::2, 0::2] + level[1::2, 1::2]
        level = (s + 2) >> 2
        y0 = PBR_TILE - 2 * nd
        band[y0:y0 + nd, PBR_TILE:PBR_TILE + nd] = level.astype(np.uint8)
        dim = nd
    return band


def atlas(bases):
    """The atlas of tile bases (each 256 x 256 x 4 uint8, or a constant RGBA
    tuple); tile 0 is the engine's reserved white band."""
    bands = [np.full((PBR_TILE, PBR_ATLAS_W, 4), 255, np.uint8)]
    for b in bases:
        if np.asarray(b).size == 4:
            b = np.broadcast_to(np.asarray(b, np.uint8), (PBR_TILE, PBR_TILE, 4))
        bands.append(atlas_band(b))
    return np.concatenate(bands, axis=0)


def batch_against_scenes(device, cases, frames=2):
    """Each (scene data, frame) of cases rendered alone (frames consecutive
    frames) and all of them as one batch: per case the two accumulations, and
    the ray counts (primary, extension, shadow) summed over the cases, alone
    and batched."""
    alone, counts = [], [0, 0, 0]
    for sd, f in cases:
        sc = dmt.Scene(device, sd)
        sc.render_frames(f, frames)
        alone.append(sc.readback(f.width, f.height))
        st = sc.stats()
        for i, k in enumerate(("primary_rays", "extension_rays", "shadow_rays")):
            counts[i] += st[k]
        sc.close()
    scenes = [dmt.Scene(device, sd) for sd, _ in cases]
    b = dmt.Batch(device, scenes, [f for _, f in cases])
    b.render(0, frames)
    st = b.stats()
    batched = [sc.readback(f.width, f.height) for sc, (_, f) in zip(scenes, cases)]
    b.close()
    for sc in scenes:
        sc.close()
    return alone, batched, tuple(counts), (st["primary_rays"], st["extension_rays"], st["shadow_rays"])
