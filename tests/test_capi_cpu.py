"""C ABI checks that need no GPU: the library loads, exports every symbol the
headers declare, and the host-only entry points validate their arguments."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import dmt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("dmt.h", "dmt_scenegen.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(dmt_[a-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    L = dmt.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_abi_version():
    v = C.c_uint32()
    assert dmt.lib().dmt_get_abi_version(C.byref(v)) == 0
    assert v.value == dmt.ABI_VERSION == 8
    assert dmt.lib().dmt_get_abi_version(None) == dmt.DMT_ERR_INVALID_ARG


def test_generators_shapes():
    sd = dmt.generate(0, seed=1)
    assert sd.prims.shape == (65, 28) and sd.lights.shape == (1, 20)
    assert sd.frame.width == 256 and sd.frame.max_bounces == 4
    sd = dmt.generate(1, seed=1, subdivisions=2, random_tris=100)
    assert sd.indices.shape == (20 * 16 + 100, 3)
    assert sd.env.shape == (256, 512, 3) and sd.frame.max_bounces == 8
    sd = dmt.generate(4, seed=1, spheres=10)
    assert sd.prims.shape == (11, 28) and sd.frame.spp == 16
    with pytest.raises(dmt.DmtError) as e:
        dmt.generate(9, seed=1)
    assert e.value.status == dmt.DMT_ERR_UNSUPPORTED


def test_full_c1_generator_counts():
    sd = dmt.generate(1, seed=1)
    assert sd.indices.shape[0] == 20 * 4 ** 8 + 200000 == 1510720
    assert sd.frame.width == 1920 and sd.frame.height == 1080
    e = sd.indices[20 * 4 ** 8:]
    tri = sd.positions[e]
    edges = np.linalg.norm(tri - np.roll(tri, 1, axis=1), axis=2)
    assert edges.max() <= 0.05 + 1e-6
    assert np.all(np.abs(tri) <= 5.0 + 0.03)


def test_host_builder_argument_errors():
    L = dmt.lib()
    cnt = C.c_uint32()
    pos = np.zeros(9, np.float32)
    bad = np.array([0, 1, 7], np.uint32)
    rc = L.dmt_build_bvh4(pos.ctypes.data_as(C.POINTER(C.c_float)), 3, bad.ctypes.data_as(C.POINTER(C.c_uint32)),
                          3, 0, 0, None, None, C.byref(cnt), None)
    assert rc == dmt.DMT_ERR_INVALID_ARG
    rc = L.dmt_build_bvh4(pos.ctypes.data_as(C.POINTER(C.c_float)), 3, bad.ctypes.data_as(C.POINTER(C.c_uint32)),
                          2, 0, 0, None, None, C.byref(cnt), None)
    assert rc == dmt.DMT_ERR_INVALID_ARG   # not a multiple of 3


def test_camera_basis_matches_reference_camera():
    """Camera.h: right = normalize(cross(fwd, +Y)), up = normalize(cross(right, fwd))."""
    f = dmt.look_at((0, 2, 12), (0, 0, 0), 50.0, 1920, 1080)
    fwd = np.array(f.fwd_aspect[:3]); r = np.array(f.right_xyz[:3]); u = np.array(f.up_xyz[:3])
    r_ref = np.cross(fwd, [0, 1, 0]); r_ref /= np.linalg.norm(r_ref)
    u_ref = np.cross(r_ref, fwd); u_ref /= np.linalg.norm(u_ref)
    assert np.allclose(r, r_ref, atol=1e-7) and np.allclose(u, u_ref, atol=1e-7)
    assert f.pos_fovtan[3] == pytest.approx(np.tan(np.radians(25.0)), rel=1e-6)
    assert f.fwd_aspect[3] == pytest.approx(1920 / 1080)


def test_device_create_reports_missing_device_without_a_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(dmt.DmtError) as e:
        dmt.Device(0)
    assert e.value.status == dmt.DMT_ERR_NO_DEVICE


def test_batch_workload_generator():
    """The batched workload's per-scene draws: resolutions from the four
    squares, triangle counts log-uniform in [10k, 2M] split into the finest
    fitting icosphere plus random triangles, spheres on odd scenes."""
    res, tris = set(), []
    for i in range(64):
        o, seed, mb = dmt.batch_workload_scene(1, i)
        assert o.width == o.height and o.width in (128, 256, 512, 1024)
        t = 20 * 4 ** o.subdivisions + o.random_tris
        assert 10_000 <= t <= 2_000_000 and o.random_tris >= 0
        if o.subdivisions < 8:
            assert 20 * 4 ** (o.subdivisions + 1) > t
        assert o.spheres == (64 if i % 2 else 0) and mb == 6
        res.add(o.width)
        tris.append(t)
    assert res == {128, 256, 512, 1024}
    assert dmt.batch_workload_scene(1, 5)[1] == dmt.batch_workload_scene(1, 5)[1]
    assert dmt.batch_workload_scene(1, 5)[1] != dmt.batch_workload_scene(2, 5)[1]
    assert 40_000 < np.median(tris) < 600_000   # log-uniform median is ~141k


def test_config1_spheres_option_keeps_the_default_stream():
    a = dmt.generate(1, seed=3, subdivisions=2, random_tris=50)
    b = dmt.generate(1, seed=3, subdivisions=2, random_tris=50, spheres=0)
    c = dmt.generate(1, seed=3, subdivisions=2, random_tris=50, spheres=64)
    assert np.array_equal(a.positions, b.positions) and np.array_equal(a.env, b.env)
    assert np.array_equal(a.positions, c.positions) and np.array_equal(a.env, c.env)
    assert a.prims.shape[0] == 0 and c.prims.shape[0] == 65
    assert c.prims[0, 8] == 1.0 and np.all(c.prims[1:, 8] == 0.0)   # ground plane, then spheres


def test_batch_entry_points_validate_without_a_device():
    h = C.c_void_p()
    assert dmt.lib().dmt_batch_create(None, None, None, 0, None, C.byref(h)) == dmt.DMT_ERR_INVALID_ARG
    assert dmt.lib().dmt_batch_render(None, 0, 1) == dmt.DMT_ERR_INVALID_ARG
    assert dmt.lib().dmt_batch_autotune(None, 1, None, None) == dmt.DMT_ERR_INVALID_ARG
    dmt.lib().dmt_batch_destroy(None)
