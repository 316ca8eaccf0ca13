"""SDF clusters through the C ABI, against the oracle bit for bit: every shape
and op in Lambert, metal and glass clusters, the many-light workload with its
boxes and unions, closest-hit queries (ties to the lower cluster index), the
frame's step cap and surface distance, the G-buffer, a batch, and the engine
path (sdf_params, the clusters at buffer slot 10); invalid clusters, a
fractal root and bad trace parameters are refused."""
import numpy as np
import pytest
import torch

import _dmt_rhi as rhi
import dmt
import oracle_ctypes as orc
import scenes
from diffrender import adjoint_numpy
from rhi_engine import EngineSim, view_proj
from scenes import (MAT_DIELECTRIC, MAT_LAMBERT, MAT_METAL, SDF_BOX, SDF_CAPSULE, SDF_DISPLACE, SDF_INTERSECT,
                    SDF_LEAF, SDF_PLANE, SDF_ROUND_BOX, SDF_SPHERE, SDF_SUBTRACT, SDF_TORUS, SDF_UNION, sdf_box,
                    sdf_cluster, sdf_node)

pytestmark = pytest.mark.gpu


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


def _kinds():
    """One cluster of every shape and op, spread over a 4 x 3 grid."""
    B = lambda c, e: (np.asarray(c) - e, np.asarray(c) + e)
    out = []
    cs = [(x, y, 0.0) for y in (1.2, 0.0, -1.2) for x in (-2.4, -0.8, 0.8, 2.4)]
    mats = [MAT_LAMBERT, MAT_METAL, MAT_DIELECTRIC]

    def add(i, nodes, ext):
        m = mats[i % 3]
        out.append(sdf_cluster(nodes, *B(cs[i], ext + 1e-3), mat=m, albedo=(0.8, 0.6, 0.4) if m != 2 else (1, 1, 1),
                               rough=0.2 if m == 1 else 0.0, ior=1.5 if m == 2 else 1.0))

    L = lambda shape, q, c: sdf_node(SDF_LEAF, shape, params=q, center=c)
    c = cs
    add(0, [L(SDF_SPHERE, (0.5, 0, 0, 0), c[0])], 0.5)
    add(1, [L(SDF_BOX, (0.45, 0.35, 0.3, 0), c[1])], 0.45)
    add(2, [L(SDF_ROUND_BOX, (0.35, 0.3, 0.25, 0.1), c[2])], 0.45)
    add(3, [L(SDF_TORUS, (0.35, 0.12, 0, 0), c[3])], 0.47)
    add(4, [L(SDF_CAPSULE, (0.3, 0.15, 0, 0), c[4])], 0.45)
    add(5, [L(SDF_BOX, (0.3, 0.3, 0.3, 0), c[5]), L(SDF_SPHERE, (0.3, 0, 0, 0), np.add(c[5], (0.25, 0.2, 0.1))),
            sdf_node(SDF_UNION, 0, 0, 1)], 0.6)
    add(6, [L(SDF_SPHERE, (0.28, 0, 0, 0), np.add(c[6], (-0.2, 0, 0))),
            L(SDF_CAPSULE, (0.2, 0.12, 0, 0), np.add(c[6], (0.22, 0, 0))),
            sdf_node(SDF_UNION, 0, 0, 1, params=(0.15, 0, 0, 0))], 0.55)
    add(7, [L(SDF_BOX, (0.4, 0.4, 0.4, 0), c[7]), L(SDF_SPHERE, (0.35, 0, 0, 0), np.add(c[7], (0.4, 0.4, 0.4))),
            sdf_node(SDF_SUBTRACT, 0, 0, 1)], 0.4)
    add(8, [L(SDF_SPHERE, (0.5, 0, 0, 0), c[8]), L(SDF_BOX, (0.35, 0.35, 0.35, 0), c[8]),
            sdf_node(SDF_INTERSECT, 0, 0, 1, params=(0.05, 0, 0, 0))], 0.5)
    add(9, [L(SDF_SPHERE, (0.35, 0, 0, 0), c[9]), sdf_node(SDF_DISPLACE, 0, 0, 0, params=(-0.1, 0, 0, 0))], 0.45)
    add(10, [L(SDF_PLANE, (0.0, 1.0, 0.0, 0.0), c[10]), L(SDF_SPHERE, (0.45, 0, 0, 0), c[10]),
             sdf_node(SDF_INTERSECT, 0, 0, 1)], 0.45)   # a half ball
    add(11, [L(SDF_BOX, (0.3, 0.3, 0.3, 0), c[11]), sdf_node(7, 0, 0, 0, params=(1.0, 0, 0, 0)),
             sdf_node(SDF_UNION, 0, 0, 1, params=(0.1, 0, 0, 0))], 0.45)   # with a procedural op (1e30)
    return out


def _kinds_scene(env=None):
    floor = scenes.plane((0, 1, 0), 2.0, albedo=(0.6, 0.6, 0.6))
    lights = [scenes.quad_light((0, 4.0, 1.0), (0, -1, 0), (1.5, 0, 0), 1.5, (12, 12, 12)),
              scenes.point_light((-3, 3, 3), (8, 7, 6))]
    sd = scenes.scene([floor, scenes.sphere((0.0, 0.0, -1.5), 0.6, albedo=(0.3, 0.5, 0.7))], lights,
                      env=env if env is not None else scenes.black_env(), sdf=_kinds())
    sd.frame = scenes.frame((0.0, 0.5, 6.0), (0.0, 0.0, 0.0), 50.0, 64, 48, spp=2, bounces=5)
    return sd


def test_every_shape_and_op_equals_the_oracle(device):
    sd = _kinds_scene()
    got, ref, counts, oc = _both(device, sd, sd.frame)
    _check("SDF shapes and ops", got, ref, counts, oc)
    bare = scenes.scene(sd.prims, sd.lights, env=sd.env)
    plain, _, _, _ = _both(device, bare, sd.frame)
    assert np.mean(np.any(plain != got, axis=-1)) > 0.2
    # with an env map (MIS, env NEE through glass clusters)
    env = np.random.default_rng(4).lognormal(0.0, 1.0, (16, 32, 3)).astype(np.float32)
    sd2 = _kinds_scene(env=env)
    got, ref, counts, oc = _both(device, sd2, sd2.frame)
    _check("SDF shapes and ops under an env map", got, ref, counts, oc)


def test_many_light_workload_with_boxes_and_unions(device):
    sd = dmt.generate(2, seed=5, spheres=400, sdf_clusters=128, subdivisions=3, random_tris=3000, lights=128,
                      width=80, height=48)
    assert sd.sdf.shape == (128, 112)
    f = sd.frame.copy()
    f.spp = 2
    got, ref, counts, oc = _both(device, sd, f)
    _check("many-light workload with SDF clusters", got, ref, counts, oc)


def test_closest_hit_queries_equal_the_oracle(device):
    cl = _kinds()
    dup = [cl[1].copy(), cl[1].copy()]   # identical clusters: the lower index wins
    dup[1][8:11] = (0.1, 0.9, 0.1)
    sd = scenes.scene([scenes.sphere((0.0, 0.0, -1.5), 0.6)], sdf=cl + dup)
    rng = np.random.default_rng(17)
    o = rng.uniform(-4, 4, (20000, 3)).astype(np.float32)
    o[:, 2] = 5.0
    tgt = rng.uniform(-3, 3, (20000, 3))
    tgt[:, 2] = rng.uniform(-0.5, 0.5, 20000)
    d = (tgt - o) / np.linalg.norm(tgt - o, axis=1, keepdims=True)
    d = d.astype(np.float32)
    sc = dmt.Scene(device, sd)
    g = sc.trace(o, d)
    sc.close()
    c = orc.OracleScene.from_scene_data(sd).trace(o, d)
    for k in ("hit", "prim", "mat", "user_id", "t", "normal"):
        assert np.array_equal(g[k], c[k]), k
    assert np.sum(g["hit"] == 3) > 1000
    assert not np.any((g["hit"] == 3) & (g["prim"] == len(cl) + 1))   # the duplicate's later copy never wins


def test_frame_step_cap_and_surface_distance(device):
    sd = _kinds_scene()
    f = sd.frame.copy()
    f.sdf_max_iters, f.sdf_epsilon = 6, 3e-3
    got, ref, _, _ = _both(device, sd, f)
    _check("6 steps, eps 3e-3", got, ref)
    full, _, _, _ = _both(device, sd, sd.frame)
    assert not np.array_equal(got, full)


def test_gbuffer_of_sdf_clusters(device):
    sd = _kinds_scene()
    f = sd.frame.copy()
    vc = view_proj(f).astype(np.float32)
    n = f.width * f.height
    bufs = {k: torch.full((n * 4,), -7.0, dtype=torch.float32, device="cuda:0") for k in ("depth", "normal", "albedo")}
    torch.cuda.synchronize()
    sc = dmt.Scene(device, sd)
    sc.gbuffer(f, vc, vc, 4, **{k: v.data_ptr() for k, v in bufs.items()})
    device.synchronize()
    sc.close()
    want = orc.OracleScene.from_scene_data(sd).gbuffer(f, vc, vc)
    got = {k: v.cpu().numpy().reshape(f.height, f.width, 4) for k, v in bufs.items()}
    assert np.array_equal(got["depth"][..., 0], want[..., 0])
    assert np.array_equal(got["normal"][..., :3], want[..., 3:6])
    assert np.array_equal(got["albedo"][..., :3], want[..., 6:9])


@pytest.mark.parametrize("plain", [0, 1], ids=["graph", "plain launches"])
def test_batch_with_an_sdf_scene_equals_per_scene_renders(device, plain):
    a = _kinds_scene()
    b = dmt.generate(0, seed=19, width=36, height=24)
    data = (a, b)
    frames = [a.frame.copy(), b.frame.copy()]
    scs = [dmt.Scene(device, sd) for sd in data]
    batch = dmt.Batch(device, scs, frames, dmt.BatchParams(plain_launches=plain))
    batch.render(0, 2)
    device.synchronize()
    for sc, sd, f in zip(scs, data, frames):
        got = sc.readback(f.width, f.height)
        one = dmt.Scene(device, sd)
        one.render_frames(f, 2)
        want = one.readback(f.width, f.height)
        one.close()
        assert np.array_equal(got, want)
    batch.close()
    for sc in scs:
        sc.close()


def test_engine_push_renders_sdf_clusters(device):
    """sdf_params (count, r_sdf_max_iters), sdf_params_f.x (r_sdf_epsilon) and
    the cluster records at buffer slot 10, as Engine.cpp:5076-5190 and
    :7947-7981 upload and push them."""
    sd = _kinds_scene()
    f = sd.frame.copy()
    f.sdf_max_iters, f.sdf_epsilon = 96, 2e-4
    dev = rhi.HipDevice(strict=True)
    eng = EngineSim(dev, sd)
    buf = dev.create_buffer(sd.sdf.nbytes)
    dev.write_buffer(buf, np.ascontiguousarray(sd.sdf, np.float32))
    ov = {"sdf_params": ("4I", (sd.sdf.shape[0], 96, 0, 12)), "sdf_params_f": ("4f", (2e-4, 8.0, 1.0, 0.0))}
    eng.frames(f, 2, buffers={10: buf}, overrides=ov)
    assert dev.last_error() == "", dev.last_error()
    got = eng.accum_image()
    dev.close()
    sc = dmt.Scene(device, sd)
    sc.render_frames(f, 2)
    want = sc.readback(f.width, f.height)
    sc.close()
    assert np.array_equal(got, want)


def test_invalid_clusters_and_parameters_are_refused(device):
    good = sdf_box((0, 0, 0), (0.3, 0.3, 0.3))
    cases = []
    bad = good.copy()
    bad.view(np.uint32)[3] = 9                       # node_count
    cases.append((bad, dmt.DMT_ERR_INVALID_ARG))
    bad = good.copy()
    bad.view(np.uint32)[7] = 4                       # material (3 is water)
    cases.append((bad, dmt.DMT_ERR_INVALID_ARG))
    bad = good.copy()
    bad[0] = 1.0                                     # aabb_min.x > aabb_max.x
    cases.append((bad, dmt.DMT_ERR_INVALID_ARG))
    bad = sdf_cluster([sdf_node(SDF_UNION, 0, 0, 1)], (-1, -1, -1), (1, 1, 1))   # children at or after the op
    cases.append((bad, dmt.DMT_ERR_INVALID_ARG))
    bad = sdf_cluster([sdf_node(SDF_LEAF, 6, params=(1, 0, 0, 0))], (-1, -1, -1), (1, 1, 1))   # a Mandelbulb root
    cases.append((bad, dmt.DMT_ERR_UNSUPPORTED))
    for cl, status in cases:
        with pytest.raises(dmt.DmtError) as e:
            dmt.Scene(device, scenes.scene(sdf=[cl]))
        assert e.value.status == status
    sd = scenes.scene([scenes.plane((0, 1, 0), 1.0)], sdf=[good])
    sc = dmt.Scene(device, sd)
    f = scenes.frame((0, 0, 3), (0, 0, 0), 40.0, 8, 8)
    for field, value in (("sdf_max_iters", 0), ("sdf_max_iters", 257), ("sdf_epsilon", 0.0),
                         ("sdf_epsilon", float("nan"))):
        g = f.copy()
        setattr(g, field, value)
        with pytest.raises(dmt.DmtError) as e:
            sc.render(g)
        assert e.value.status == dmt.DMT_ERR_INVALID_ARG, field
    g = adjoint_numpy(sc, f, np.ones((8, 8, 4), np.float32), device)   # the adjoint runs on cluster scenes
    assert np.all(np.isfinite(g))
    sc.render(f)   # the scene still renders
    sc.close()
    assert not device.is_lost()
