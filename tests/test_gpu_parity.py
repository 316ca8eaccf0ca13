"""Parity of the HIP path against the CPU oracle, through the C ABI.

Bars (from the workload definition):
  * first-hit primitive / material / user ids bit-exact;
  * hit distance and normal within 1e-5 relative (the implementation is
    bit-exact by construction; the tests report the exact-match fraction too);
  * radiance at equal spp within relative MSE 1e-4 (same PCG stream).
"""
import numpy as np
import pytest

import dmt
import oracle_ctypes as orc
from conftest import rel_mse

pytestmark = pytest.mark.gpu

T_REL = 1e-5
RADIANCE_REL_MSE = 1e-4


def _random_rays(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


def _camera_rays(frame, step=7):
    W, H = frame.width, frame.height
    ys, xs = np.mgrid[0:H:step, 0:W:step]
    u = ((xs + 0.5) / W) * 2 - 1
    v = -(((ys + 0.5) / H) * 2 - 1)
    f = np.array(frame.fwd_aspect[:3]); r = np.array(frame.right_xyz[:3]); up = np.array(frame.up_xyz[:3])
    t, a = frame.pos_fovtan[3], frame.fwd_aspect[3]
    d = f[None] + r[None] * (u.reshape(-1, 1) * a * t) + up[None] * (v.reshape(-1, 1) * t)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.tile(np.array(frame.pos_fovtan[:3], np.float32), (d.shape[0], 1))
    return o, d.astype(np.float32)


def _compare_hits(g, c):
    assert np.array_equal(g["hit"], c["hit"]), f"hit kind differs on {(g['hit'] != c['hit']).sum()} rays"
    hit = c["hit"] > 0
    assert hit.any()
    assert np.array_equal(g["prim"][hit], c["prim"][hit]), "primitive ids differ"
    assert np.array_equal(g["mat"][hit], c["mat"][hit]), "material ids differ"
    assert np.array_equal(g["user_id"][hit], c["user_id"][hit]), "user ids differ"
    tg, tc = g["t"][hit].astype(np.float64), c["t"][hit].astype(np.float64)
    assert np.all(np.abs(tg - tc) <= T_REL * np.abs(tc)), "hit distance beyond 1e-5 relative"
    ng, nc = g["normal"][hit].astype(np.float64), c["normal"][hit].astype(np.float64)
    assert np.all(np.abs(ng - nc) <= T_REL * np.maximum(np.abs(nc), 1e-3)), "normal beyond 1e-5"
    return float(np.mean(tg == tc)), float(np.mean(np.all(ng == nc, axis=1)))


@pytest.mark.parametrize("cfg,kw", [(0, {}), (1, dict(subdivisions=5, random_tris=20000))])
def test_trace_closest_matches_oracle(device, cfg, kw):
    sd = dmt.generate(cfg, seed=3, **kw)
    scene = dmt.Scene(device, sd)
    oscene = orc.OracleScene.from_scene_data(sd)
    for o, d in (_random_rays(20000, -6, 6, 11), _camera_rays(sd.frame)):
        g = scene.trace(o, d)
        c = oscene.trace(o, d)
        t_exact, n_exact = _compare_hits(g, c)
        print(f"config {cfg}: {len(o)} rays, hit {np.mean(c['hit'] > 0):.3f}, t exact {t_exact:.6f}, "
              f"normal exact {n_exact:.6f}")
        assert t_exact == 1.0 and n_exact == 1.0


def _render_pair(device, sd, frame, frames=1):
    scene = dmt.Scene(device, sd)
    oscene = orc.OracleScene.from_scene_data(sd)
    acc_o = None
    for k in range(frames):
        f = frame.copy()
        f.frame_index = frame.frame_index + k
        f.reset_accum = 1 if k == 0 else 0
        scene.render(f)
        acc_o, cnt = oscene.render(f, 32, 0, accum=acc_o)
    img = scene.readback(frame.width, frame.height)
    st = scene.stats()
    return img, acc_o, st, cnt


def _check_render(img, ref, st, cnt, label):
    err = rel_mse(img[..., :3], ref[..., :3])
    exact = float(np.mean(np.all(img == ref, axis=-1)))
    print(f"{label}: rel-MSE {err:.3e}, bit-exact pixels {exact:.6f}, gpu rays "
          f"{st['primary_rays']}/{st['extension_rays']}/{st['shadow_rays']}, oracle {cnt}")
    assert err <= RADIANCE_REL_MSE
    assert np.array_equal(img[..., 3], ref[..., 3])
    return exact


def test_render_c0_matches_oracle(device):
    sd = dmt.generate(0, seed=1)
    img, ref, st, cnt = _render_pair(device, sd, sd.frame)
    exact = _check_render(img, ref, st, cnt, "C0 256x256 1spp d4")
    assert (st["primary_rays"], st["extension_rays"], st["shadow_rays"]) == \
        (cnt["primary"], cnt["extension"], cnt["shadow"])
    assert exact > 0.999


def test_render_c1_small_matches_oracle(device):
    sd = dmt.generate(1, seed=2, subdivisions=5, random_tris=30000, width=320, height=180)
    img, ref, st, cnt = _render_pair(device, sd, sd.frame)
    exact = _check_render(img, ref, st, cnt, "C1-small 320x180 1spp d8")
    assert (st["primary_rays"], st["extension_rays"], st["shadow_rays"]) == \
        (cnt["primary"], cnt["extension"], cnt["shadow"])
    assert exact > 0.999


def test_render_spp4_accumulate_matches_oracle(device):
    sd = dmt.generate(4, seed=5, spheres=48, width=96, height=64)
    f = sd.frame.copy()
    f.spp = 4
    img, ref, st, cnt = _render_pair(device, sd, f, frames=3)
    _check_render(img, ref, st, cnt, "C4-small 96x64 4spp x3 frames")
    assert np.all(img[..., 3] == 3.0)


def test_render_no_env_matches_oracle(device):
    sd = dmt.generate(0, seed=4, width=128, height=96)
    sd.env = None
    img, ref, st, cnt = _render_pair(device, sd, sd.frame)
    _check_render(img, ref, st, cnt, "C0 without env map")


def test_render_c1_full_size_rows(device):
    """Full C1 (1.51 M triangles, 1920x1080, depth 8) on the GPU; the oracle
    renders every 45th row (24 rows) and those rows must match."""
    sd = dmt.generate(1, seed=1)
    scene = dmt.Scene(device, sd)
    scene.render(sd.frame)
    img = scene.readback(sd.frame.width, sd.frame.height)
    st = scene.stats()
    oscene = orc.OracleScene.from_scene_data(sd)
    ref, cnt = oscene.render(sd.frame, 32, 0, rows=(0, sd.frame.height, 45))
    rows = np.arange(0, sd.frame.height, 45)
    err = rel_mse(img[rows, :, :3], ref[rows, :, :3])
    exact = float(np.mean(np.all(img[rows] == ref[rows], axis=-1)))
    print(f"C1 full: rel-MSE on {len(rows)} rows {err:.3e}, bit-exact {exact:.6f}, rays {st}")
    assert err <= RADIANCE_REL_MSE
    assert st["primary_rays"] == sd.frame.width * sd.frame.height


def test_render_c2_small_matches_oracle(device):
    """Config 2 at reduced size: spheres + mesh lit only by emissive
    triangles through a 256-leaf light tree, 2 spp, depth 6."""
    sd = dmt.generate(2, seed=3, spheres=300, subdivisions=4, random_tris=20000, lights=256, width=160, height=90)
    f = sd.frame.copy()
    f.spp = 2
    img, ref, st, cnt = _render_pair(device, sd, f)
    exact = _check_render(img, ref, st, cnt, "C2-small 160x90 2spp d6, 256 triangle lights")
    assert (st["primary_rays"], st["extension_rays"], st["shadow_rays"]) == \
        (cnt["primary"], cnt["extension"], cnt["shadow"])
    assert exact > 0.999


def test_render_c2_full_rows(device):
    """Full config 2 (4,096 spheres, 500k triangles, 4,096 triangle lights,
    1280x720, 4 spp, depth 6) on the GPU; the oracle renders every 60th row."""
    sd = dmt.generate(2, seed=1)
    scene = dmt.Scene(device, sd)
    scene.render(sd.frame)
    img = scene.readback(sd.frame.width, sd.frame.height)
    oscene = orc.OracleScene.from_scene_data(sd)
    ref, cnt = oscene.render(sd.frame, 32, 0, rows=(0, sd.frame.height, 60))
    rows = np.arange(0, sd.frame.height, 60)
    err = rel_mse(img[rows, :, :3], ref[rows, :, :3])
    exact = float(np.mean(np.all(img[rows] == ref[rows], axis=-1)))
    print(f"C2 full: rel-MSE on {len(rows)} rows {err:.3e}, bit-exact {exact:.6f}")
    assert err <= RADIANCE_REL_MSE
    assert exact > 0.999


@pytest.mark.parametrize("cfg,kw,spp", [(0, dict(seed=6, width=128, height=96), 1),
                                        (4, dict(seed=7, spheres=40, width=64, height=48), 4),
                                        (1, dict(seed=8, subdivisions=4, random_tris=5000, width=160, height=90), 1)])
def test_render_frames_equals_sequential_frames(device, cfg, kw, spp):
    """dmt_render_frames(K) regenerates paths across frames inside one call;
    its image must be bit-identical to K dmt_render calls and to the oracle."""
    sd = dmt.generate(cfg, **kw)
    f = sd.frame.copy()
    f.spp = spp
    f.frame_index = 3
    K = 5
    a = dmt.Scene(device, sd)
    a.render_frames(f, K)
    img_a = a.readback(f.width, f.height)
    st_a = a.stats()
    b = dmt.Scene(device, sd)
    ref = None
    oscene = orc.OracleScene.from_scene_data(sd)
    cnt_total = {"primary": 0, "extension": 0, "shadow": 0}
    for k in range(K):
        fk = f.copy()
        fk.frame_index = f.frame_index + k
        fk.reset_accum = 1 if k == 0 else 0
        b.render(fk)
        ref, cnt = oscene.render(fk, 32, 0, accum=ref)
        for key in cnt_total:
            cnt_total[key] += cnt[key]
    img_b = b.readback(f.width, f.height)
    print(f"config {cfg}: streaming vs per-frame equal {np.array_equal(img_a, img_b)}, "
          f"vs oracle rel-MSE {rel_mse(img_a[..., :3], ref[..., :3]):.3e}")
    assert np.array_equal(img_a, img_b)
    assert np.array_equal(img_a, ref)
    assert (st_a["primary_rays"], st_a["extension_rays"], st_a["shadow_rays"]) == \
        (cnt_total["primary"], cnt_total["extension"], cnt_total["shadow"])
    assert np.all(img_a[..., 3] == K)


_SCHEDULE_SCENES = [
    ("C0 1spp", 0, dict(seed=6, width=128, height=96), 1, False),
    ("C0 no env", 0, dict(seed=4, width=96, height=64), 1, True),
    ("C4 4spp", 4, dict(seed=7, spheres=40, width=64, height=48), 4, False),
    ("C1 small", 1, dict(seed=8, subdivisions=4, random_tris=5000, width=160, height=90), 1, False),
    ("C2 small 2spp", 2, dict(seed=3, spheres=200, subdivisions=3, random_tris=5000, lights=64, width=96,
                              height=54), 2, False),
]


@pytest.mark.parametrize("label,cfg,kw,spp,no_env", _SCHEDULE_SCENES, ids=[s[0] for s in _SCHEDULE_SCENES])
def test_schedules_render_identical_images(device, label, cfg, kw, spp, no_env):
    """The frame-completion launch (k_tail) after any number of per-bounce
    iterations gives the image, the ray counts and the traversal totals of
    per-bounce kernels to the end, over three accumulated frames."""
    sd = dmt.generate(cfg, **kw)
    if no_env:
        sd.env = None
    f = sd.frame.copy()
    f.spp = spp
    scene = dmt.Scene(device, sd)
    results = {}
    for sched in (dmt.SCHEDULE_WAVEFRONT, 0, 1, 2, 3, 5):
        scene.set_schedule(sched)
        assert scene.schedule() == sched
        scene.reset_stats()
        for k in range(3):
            fk = f.copy()
            fk.frame_index = 11 + k
            fk.reset_accum = 1 if k == 0 else 0
            scene.render(fk)
        img = scene.readback(f.width, f.height)
        st = scene.stats()
        tr = scene.measure_traffic(f)
        results[sched] = (img, (st["primary_rays"], st["extension_rays"], st["shadow_rays"]),
                          (tr["node_visits"], tr["tri_tests"], tr["sphere_tests"]))
    ref_img, ref_rays, ref_trav = results[dmt.SCHEDULE_WAVEFRONT]
    assert ref_rays[0] == 3 * spp * f.width * f.height
    for sched, (img, rays, trav) in results.items():
        print(f"{label}: schedule {sched}: image equal {np.array_equal(img, ref_img)}, rays {rays}, traversal {trav}")
        assert np.array_equal(img, ref_img), f"schedule {sched}"
        assert rays == ref_rays, f"schedule {sched}"
        assert trav == ref_trav, f"schedule {sched}"
    oscene = orc.OracleScene.from_scene_data(sd)
    ref = None
    for k in range(3):
        fk = f.copy()
        fk.frame_index = 11 + k
        fk.reset_accum = 1 if k == 0 else 0
        ref, _ = oscene.render(fk, 32, 0, accum=ref)
    assert rel_mse(ref_img[..., :3], ref[..., :3]) <= RADIANCE_REL_MSE


def test_error_paths_through_the_abi(device):
    sd = dmt.generate(0, seed=1, width=16, height=16)
    sc = dmt.Scene(device, sd)
    f = sd.frame.copy()
    f.spp = 0
    with pytest.raises(dmt.DmtError) as e:
        sc.render(f)
    assert e.value.status == dmt.DMT_ERR_INVALID_ARG
    f = sd.frame.copy()
    f.max_bounces = 65
    with pytest.raises(dmt.DmtError) as e:
        sc.render(f)
    assert e.value.status == dmt.DMT_ERR_LIMIT
    with pytest.raises(dmt.DmtError):
        sc.render_frames(sd.frame, 0)
    bad = dmt.generate(0, seed=1)
    bad.prims[3, 9] = 7.0   # material id outside 0..2
    with pytest.raises(dmt.DmtError) as e:
        dmt.Scene(device, bad)
    assert e.value.status == dmt.DMT_ERR_INVALID_ARG
    bad = dmt.generate(1, seed=1, subdivisions=1, random_tris=0)
    bad.indices[2, 1] = 10_000
    with pytest.raises(dmt.DmtError):
        dmt.Scene(device, bad)
    # readback size must match exactly (full-or-nothing, SoftwareDevice.cpp:458-483)
    sc.render(sd.frame)
    out = np.zeros(10, np.float32)
    import ctypes as C
    rc = dmt.lib().dmt_readback_accum(sc._h, out.ctypes.data_as(C.POINTER(C.c_float)), out.size)
    assert rc == dmt.DMT_ERR_INVALID_ARG
    assert not device.is_lost()
