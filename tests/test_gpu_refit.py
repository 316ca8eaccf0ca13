"""Moving meshes: dmt_scene_update_mesh refits the triangle BVH on the device.
The closest hit does not depend on the tree, so a refit scene renders exactly
the images of a scene built from the new positions (and of the oracle), for
small deformations and for large rigid motions; trace queries match the
oracle's exhaustive scan; and through the RHI device a new TLAS over the same
triangles takes the refit path."""
import time

import numpy as np
import pytest

import _dmt_rhi as rhi
import dmt
import oracle_ctypes as orc
from rhi_engine import EngineSim

pytestmark = pytest.mark.gpu


def _mesh_scene(seed=51, w=64, h=40):
    return dmt.generate(1, seed=seed, subdivisions=4, random_tris=3000, width=w, height=h)


def _with_positions(sd, pos):
    return dmt.SceneData(prims=sd.prims, lights=sd.lights, positions=pos, indices=sd.indices, env=sd.env,
                         env_intensity=sd.env_intensity, frame=sd.frame)


def _render(device, sc, frame, k):
    f = frame.copy()
    f.reset_accum = 1
    sc.render_frames(f, k)
    return sc.readback(f.width, f.height)


def _deformations(p):
    yield "unchanged", p.copy()
    yield "smooth wave", (p + 0.15 * np.sin(3.0 * p[:, [1, 2, 0]])).astype(np.float32)
    yield "rigid move", (p * np.float32(1.3) + np.float32([2.0, -1.0, 0.5])).astype(np.float32)
    rng = np.random.default_rng(3)
    yield "jitter", (p + rng.normal(0.0, 0.05, p.shape)).astype(np.float32)


@pytest.mark.parametrize("case", ["unchanged", "smooth wave", "rigid move", "jitter"])
def test_refit_renders_the_rebuilt_scene(device, case):
    sd = _mesh_scene()
    pos = dict(_deformations(sd.positions))[case]
    sc = dmt.Scene(device, sd)
    _render(device, sc, sd.frame, 1)   # frames of the old mesh first
    sc.update_mesh(pos)
    got = _render(device, sc, sd.frame, 2)
    fresh = dmt.Scene(device, _with_positions(sd, pos))
    want = _render(device, fresh, sd.frame, 2)
    same = float(np.mean(np.all(got == want, axis=-1)))
    print(f"{case}: refit == rebuilt on {same:.6f} of pixels")
    assert np.array_equal(got, want)
    # the rays the two trees traverse differ, the hits do not
    st = sc.measure_traffic(sd.frame)
    st2 = fresh.measure_traffic(sd.frame)
    print(f"{case}: node visits refit {st['node_visits']} vs rebuilt {st2['node_visits']}")
    sc.close()
    fresh.close()


def test_refit_matches_the_oracle(device):
    sd = _mesh_scene(seed=52, w=40, h=30)
    pos = dict(_deformations(sd.positions))["smooth wave"]
    sc = dmt.Scene(device, sd)
    sc.update_mesh(pos)
    f = sd.frame.copy()
    f.reset_accum = 1
    sc.render(f)
    got = sc.readback(f.width, f.height)
    ref, _ = orc.OracleScene.from_scene_data(_with_positions(sd, pos)).render(f, 32, 0)
    assert np.array_equal(got, ref)
    # trace queries against the oracle's exhaustive scan
    rng = np.random.default_rng(9)
    o = rng.uniform(-6, 6, (4000, 3)).astype(np.float32)
    d = rng.normal(size=(4000, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    h = sc.trace(o, d)
    b = orc.OracleScene.from_scene_data(_with_positions(sd, pos)).trace(o, d, brute=True)
    for k in ("t", "hit", "prim"):
        assert np.array_equal(h[k], b[k]), k
    sc.close()


def test_repeated_refits_and_errors(device):
    sd = _mesh_scene(seed=53)
    sc = dmt.Scene(device, sd)
    p = sd.positions
    for k in range(4):   # a mesh moving frame after frame
        q = (p + np.float32(0.1 * k)).astype(np.float32)
        sc.update_mesh(q)
        f = sd.frame.copy()
        f.frame_index = k
        f.reset_accum = 1
        sc.render(f)
    got = sc.readback(sd.frame.width, sd.frame.height)
    fresh = dmt.Scene(device, _with_positions(sd, q))
    f = sd.frame.copy()
    f.frame_index = 3
    f.reset_accum = 1
    fresh.render(f)
    assert np.array_equal(got, fresh.readback(sd.frame.width, sd.frame.height))
    with pytest.raises(dmt.DmtError):
        sc.update_mesh(p[:-1])
    bad = p.copy()
    bad[5, 1] = np.nan
    with pytest.raises(dmt.DmtError):
        sc.update_mesh(bad)
    sp = dmt.Scene(device, dmt.generate(0, seed=1, width=16, height=16))
    with pytest.raises(dmt.DmtError):
        sp.update_mesh(p)
    sp.close()
    fresh.close()
    sc.close()


def test_refit_time_on_the_large_mesh(device):
    sd = dmt.generate(1, seed=1)          # 1.51 M triangles
    t0 = time.perf_counter()
    sc = dmt.Scene(device, sd)
    build = time.perf_counter() - t0
    pos = (sd.positions + np.float32(0.01)).astype(np.float32)
    sc.update_mesh(pos)                   # first call allocates
    t0 = time.perf_counter()
    for _ in range(5):
        sc.update_mesh(pos)
    refit = (time.perf_counter() - t0) / 5
    print(f"1.51 M triangles: host build + upload {build * 1e3:.0f} ms, device refit {refit * 1e3:.2f} ms "
          f"(includes the {pos.nbytes / 1e6:.1f} MB position upload)")
    assert refit < build
    sc.close()


def test_engine_tlas_with_moved_instances_refits(device):
    sd = _mesh_scene(seed=54)
    dev = rhi.HipDevice()
    eng = EngineSim(dev, dmt.SceneData(prims=sd.prims, lights=sd.lights, env=sd.env, env_intensity=sd.env_intensity,
                                       frame=sd.frame))
    blas = dev.create_blas(sd.positions, sd.indices)
    m1 = [1, 0, 0, 0.0, 0, 1, 0, 0.0, 0, 0, 1, 0.0]
    m2 = [1, 0, 0, 0.4, 0, 1, 0, -0.2, 0, 0, 1, 0.1]
    eng.tlas = dev.create_tlas([(blas, m1, 0, 0xFF), (blas, m2, 1, 0xFF)])
    f = sd.frame.copy()
    f.reset_accum = 1
    eng.frame(f)
    assert dev.scene_builds() == 1
    m2b = [0.9, 0.1, 0, 0.7, -0.1, 0.9, 0, -0.3, 0, 0, 1, 0.3]
    old = eng.tlas
    eng.tlas = dev.create_tlas([(blas, m1, 0, 0xFF), (blas, m2b, 1, 0xFF)])
    dev.destroy_accel(old)
    g = f.copy()
    g.frame_index += 1
    eng.frames(g, 2)
    assert dev.scene_builds() == 1 and dev.mesh_refits() == 1, dev.last_error()
    got = eng.accum_image()
    p = sd.positions
    m = np.array(m2b, np.float32)
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    t = np.stack([((m[0] * x + m[1] * y) + m[2] * z) + m[3], ((m[4] * x + m[5] * y) + m[6] * z) + m[7],
                  ((m[8] * x + m[9] * y) + m[10] * z) + m[11]], axis=1).astype(np.float32)
    flat = dmt.SceneData(prims=sd.prims, lights=sd.lights, positions=np.concatenate([p, t]),
                         indices=np.concatenate([sd.indices, sd.indices + np.uint32(p.shape[0])]), env=sd.env,
                         env_intensity=sd.env_intensity, frame=sd.frame)
    sc = dmt.Scene(device, flat)
    sc.render_frames(g, 2)
    assert np.array_equal(got, sc.readback(g.width, g.height))
    sc.close()
    dev.close()


def _device_build_cases():
    yield "mesh + env", _mesh_scene(seed=61)
    yield "spheres + mesh + lights", dmt.generate(2, seed=62, spheres=120, subdivisions=3, random_tris=2500, lights=40,
                                                  width=48, height=32)
    sd = _mesh_scene(seed=63, w=32, h=24)
    yield "one triangle", dmt.SceneData(prims=sd.prims, lights=sd.lights, positions=sd.positions,
                                        indices=sd.indices[:1], env=sd.env, frame=sd.frame)
    yield "three triangles", dmt.SceneData(prims=sd.prims, lights=sd.lights, positions=sd.positions,
                                           indices=sd.indices[100:103], env=sd.env, frame=sd.frame)
    dup = np.repeat(sd.indices[:40], 50, axis=0)      # equal Morton codes everywhere
    yield "duplicated triangles", dmt.SceneData(prims=sd.prims, lights=sd.lights, positions=sd.positions, indices=dup,
                                                env=sd.env, frame=sd.frame)


DEVICE_CASES = list(_device_build_cases())


@pytest.mark.parametrize("idx", range(len(DEVICE_CASES)), ids=[c[0] for c in DEVICE_CASES])
def test_device_build_renders_the_host_built_images(device, idx):
    label, sd = DEVICE_CASES[idx]
    host = dmt.Scene(device, sd)
    dev = dmt.Scene(device, sd, device_build=True)
    a = _render(device, host, sd.frame, 2)
    b = _render(device, dev, sd.frame, 2)
    assert np.array_equal(a, b), label
    rng = np.random.default_rng(idx)
    o = rng.uniform(-6, 6, (3000, 3)).astype(np.float32)
    d = rng.normal(size=(3000, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    h = dev.trace(o, d)
    ref = orc.OracleScene.from_scene_data(sd).trace(o, d, brute=True)
    for k in ("t", "hit", "prim"):
        assert np.array_equal(h[k], ref[k]), (label, k)
    sh, sd_ = host.stats(), dev.stats()
    print(f"{label}: {sd.indices.shape[0]} triangles, nodes host {sh['bvh_nodes']} / device {sd_['bvh_nodes']}, "
          f"depth {sh['bvh_depth']} / {sd_['bvh_depth']}")
    host.close()
    dev.close()


def test_device_build_then_refit(device):
    sd = _mesh_scene(seed=64)
    sc = dmt.Scene(device, sd, device_build=True)
    pos = dict(_deformations(sd.positions))["smooth wave"]
    sc.update_mesh(pos)
    got = _render(device, sc, sd.frame, 1)
    fresh = dmt.Scene(device, _with_positions(sd, pos))
    assert np.array_equal(got, _render(device, fresh, sd.frame, 1))
    sc.close()
    fresh.close()


def test_device_build_time_and_traversal_on_the_large_mesh(device):
    sd = dmt.generate(1, seed=1)          # 1.51 M triangles, 1920x1080
    out = {}
    for label, kw in (("host SAH", {}), ("device LBVH", {"device_build": True})):
        t0 = time.perf_counter()
        sc = dmt.Scene(device, sd, **kw)
        build = time.perf_counter() - t0
        f = sd.frame.copy()
        f.reset_accum = 1
        sc.render_frames(f, 2)
        device.synchronize()
        t0 = time.perf_counter()
        g = sd.frame.copy()
        g.frame_index = 2
        sc.render_frames(g, 10)
        device.synchronize()
        dt = (time.perf_counter() - t0) / 10
        tr = sc.measure_traffic(f)
        st = sc.stats()
        out[label] = sc.readback(f.width, f.height)
        print(f"{label}: scene build {build * 1e3:.0f} ms, {st['bvh_nodes']} nodes depth {st['bvh_depth']}, "
              f"{dt * 1e3:.3f} ms/frame, {tr['node_visits']} node visits")
        sc.close()
    assert np.array_equal(out["host SAH"], out["device LBVH"])
