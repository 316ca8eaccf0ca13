"""Batched persistent form: many scenes in one persistent-kernel launch per
frame (dmt_batch_*), through the C ABI.

The bar is identical images: every scene of a batch must be bit-identical to
the same scene rendered on its own by the wavefront path (dmt_render_frames)
and to the CPU oracle, whatever the tile, queue, stack and occupancy
parameters, with graph replay or plain launches.
"""
import numpy as np
import pytest

import dmt
import oracle_ctypes as orc

pytestmark = pytest.mark.gpu


def _scenes():
    """Small scenes of every kind: ragged sizes, spp > 1, no env map, a scene
    with max_bounces 0, and two scenes of the batched workload generator."""
    out = []
    sd = dmt.generate(0, seed=6, width=96, height=64)
    out.append(("C0 96x64", sd, sd.frame.copy()))
    sd = dmt.generate(1, seed=8, subdivisions=4, random_tris=5000, width=100, height=37)
    f = sd.frame.copy()
    f.max_bounces = 6
    out.append(("C1 100x37 d6", sd, f))
    sd = dmt.generate(4, seed=7, spheres=40, width=64, height=48)
    f = sd.frame.copy()
    f.spp = 4
    out.append(("C4 64x48 4spp", sd, f))
    sd = dmt.generate(0, seed=4, width=72, height=40)
    sd.env = None
    out.append(("C0 no env 72x40", sd, sd.frame.copy()))
    for i in (1, 2):
        sd = dmt.generate_batch_scene(5, i, scale=0.125)
        out.append((f"batch-workload #{i} {sd.frame.width}x{sd.frame.height}", sd, sd.frame.copy()))
    sd = dmt.generate(0, seed=9, width=32, height=24)
    f = sd.frame.copy()
    f.max_bounces = 0
    out.append(("max_bounces 0", sd, f))
    return out


def _with_index(frames, first_index):
    fs = []
    for k, f in enumerate(frames):
        g = f.copy()
        g.frame_index = first_index + k
        g.reset_accum = 1
        fs.append(g)
    return fs


@pytest.fixture(scope="module")
def cases():
    return _scenes()


@pytest.fixture(scope="module")
def reference(device, cases):
    """Per-scene wavefront renders (3 frames) and the oracle's images."""
    K = 3
    frames = _with_index([f for _, _, f in cases], 5)
    imgs, refs, counts = [], [], []
    for (label, sd, _), f in zip(cases, frames):
        sc = dmt.Scene(device, sd)
        sc.render_frames(f, K)
        imgs.append(sc.readback(f.width, f.height))
        st = sc.stats()
        counts.append((st["primary_rays"], st["extension_rays"], st["shadow_rays"]))
        sc.close()
        osc = orc.OracleScene.from_scene_data(sd)
        acc = None
        for k in range(K):
            fk = f.copy()
            fk.frame_index = f.frame_index + k
            fk.reset_accum = 1 if k == 0 else 0
            if f.max_bounces == 0:
                continue
            acc, _ = osc.render(fk, 32, 0, accum=acc)
        refs.append(acc)
    return frames, imgs, refs, counts, K


def test_per_scene_renders_match_oracle(reference, cases):
    frames, imgs, refs, _, K = reference
    for (label, _, _), img, ref in zip(cases, imgs, refs):
        if ref is None:
            assert np.all(img == 0.0), label   # no frame touches the buffer
            continue
        assert np.array_equal(img, ref), label


PARAMS = [
    dict(),
    dict(tile_w=64, tiles_per_fetch=4, blocks_per_cu=1, stack_lds=32),
    dict(tile_w=16, tiles_per_fetch=2, blocks_per_cu=2, stack_lds=24, plain_launches=1),
    dict(tile_w=32, tiles_per_fetch=16, blocks_per_cu=3, stack_lds=16),
]


@pytest.mark.parametrize("kw", PARAMS, ids=["default", "t64f4b1s32", "t16f2b2s24plain", "t32f16b3s16"])
def test_batch_equals_per_scene_renders(device, cases, reference, kw):
    frames, imgs, _, counts, K = reference
    scenes = [dmt.Scene(device, sd) for _, sd, _ in cases]
    p = dmt.BatchParams(**kw)
    b = dmt.Batch(device, scenes, frames, p)
    b.render(0, K)
    st = b.stats()
    for (label, _, _), sc, f, img in zip(cases, scenes, frames, imgs):
        got = sc.readback(f.width, f.height)
        same = float(np.mean(np.all(got == img, axis=-1)))
        print(f"{label}: batch == per-scene on {same:.6f} of pixels")
        assert np.array_equal(got, img), label
    want = tuple(sum(c[i] for c in counts) for i in range(3))
    assert (st["primary_rays"], st["extension_rays"], st["shadow_rays"]) == want
    assert st["frames"] == K
    b.close()
    for sc in scenes:
        sc.close()


def test_graph_replays_between_host_work(device, cases, reference):
    """One captured frame replayed frame after frame, with readbacks and
    host synchronisation between the replays; every replay must equal the
    per-scene render of the same frame count."""
    frames, _, _, _, _ = reference
    scenes = [dmt.Scene(device, sd) for _, sd, _ in cases]
    solo = [dmt.Scene(device, sd) for _, sd, _ in cases]
    b = dmt.Batch(device, scenes, frames)
    for k in range(4):
        b.render(k, 1)
        device.synchronize()
        for (label, _, _), sc, so, f in zip(cases, scenes, solo, frames):
            fk = f.copy()
            fk.frame_index = f.frame_index + k
            fk.reset_accum = 1 if k == 0 else 0
            so.render(fk)
            a = sc.readback(f.width, f.height)
            e = so.readback(f.width, f.height)
            assert np.array_equal(a, e), f"{label}: replay {k}"
    b.close()
    for sc in scenes + solo:
        sc.close()


def test_autotune_keeps_images_and_picks_valid_params(device, cases, reference):
    frames, imgs, _, _, K = reference
    scenes = [dmt.Scene(device, sd) for _, sd, _ in cases]
    b = dmt.Batch(device, scenes, frames)
    b.render(0, K)
    before = [sc.readback(f.width, f.height) for sc, f in zip(scenes, frames)]
    p, ms = b.autotune(frames_per_trial=1)
    print(f"autotuned {p.as_dict()} at {ms:.3f} ms/frame")
    assert p.tile_w in (8, 16, 32, 64) and p.stack_lds in (16, 24, 32)
    assert 1 <= p.blocks_per_cu <= 5 and 1 <= p.tiles_per_fetch <= 16 and ms > 0.0
    assert b.params().as_dict() == p.as_dict()
    after = [sc.readback(f.width, f.height) for sc, f in zip(scenes, frames)]
    for x, y in zip(before, after):
        assert np.array_equal(x, y)   # trials render into scratch buffers
    b.render(0, K)                    # again from frame 0 with the tuned parameters
    for sc, f, img in zip(scenes, frames, imgs):
        assert np.array_equal(sc.readback(f.width, f.height), img)
    b.close()
    for sc in scenes:
        sc.close()


def test_batch_error_paths(device, cases):
    _, sd, f = cases[0]
    _, sd2, f2 = cases[1]
    a, c = dmt.Scene(device, sd), dmt.Scene(device, sd2)
    with pytest.raises(dmt.DmtError) as e:
        dmt.Batch(device, [a, a], [f, f])
    assert e.value.status == dmt.DMT_ERR_INVALID_ARG
    for bad in (dict(tile_w=12), dict(stack_lds=20), dict(blocks_per_cu=6), dict(tiles_per_fetch=17)):
        with pytest.raises(dmt.DmtError) as e:
            dmt.Batch(device, [a, c], [f, f2], dmt.BatchParams(**bad))
        assert e.value.status == dmt.DMT_ERR_INVALID_ARG
    b = dmt.Batch(device, [a, c], [f, f2])
    with pytest.raises(dmt.DmtError):
        b.set_params(dmt.BatchParams(tile_w=7))
    with pytest.raises(dmt.DmtError):
        b.render(0, 0)
    # a member keeps its resolution while the batch lives
    g = f.copy()
    g.width, g.height = f.width + 8, f.height
    with pytest.raises(dmt.DmtError) as e:
        a.render(g)
    assert e.value.status == dmt.DMT_ERR_INVALID_ARG
    # destroying a member is deferred to the batch's destruction
    c.close()
    b.render(0, 2)
    img = a.readback(f.width, f.height)
    assert np.all(img[..., 3] == 2.0)
    b.close()
    a.render(g)   # free to resize now
    assert not device.is_lost()
    a.close()
