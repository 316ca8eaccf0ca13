"""GPU parity on the paths the configuration workloads leave out, through the
C ABI, against the oracle bit for bit:

  * every analytic light type of the reference (point, spot, sphere, quad)
    and the emissive triangle, alone and mixed in one light tree, with the
    tree on and off (PathTrace.slang:3979-4179);
  * EMA accumulation (r_accum_ema_alpha, PathTrace.slang:7761-7781) in the
    frame-completion kernel and in the batched kernel;
  * the sanitising step (PathTrace.slang:7722-7753): a frame value that
    overflows to inf, negative frame values, and a poisoned (NaN / inf)
    previous accumulation that restarts the pixel;
  * the full batched workload (64 scenes) equal to the per-scene renders;
  * the full differentiable workload (256 spheres, 512x512, 16 spp):
    albedo and emission gradients against central finite differences of the
    GPU forward (h = 1e-3) within 1e-3 relative, on the parameters whose
    gradient is at least 1 % of the largest;
  * convergence against the fp64 oracle: a high-spp GPU mean and an fp64
    mean over independent frames differ by Monte Carlo noise only (their
    squared difference within 1.3x of the noise the per-frame spread
    predicts, the image means within 5 sigma).
"""
import math

import numpy as np
import pytest
import torch

import dmt
import oracle_ctypes as orc
import scenes

pytestmark = pytest.mark.gpu


def _both(device, sd, f, frames=2, accum0=None):
    """GPU accumulation after `frames` frames, and the oracle's."""
    sc = dmt.Scene(device, sd)
    W, H = f.width, f.height
    buf = None
    if accum0 is not None:
        buf = torch.from_numpy(np.ascontiguousarray(accum0, np.float32)).cuda()
        torch.cuda.synchronize()
        sc.bind_accum(buf.data_ptr(), W, H)
    sc.render_frames(f, frames)
    got = sc.readback(W, H)
    st = sc.stats()
    sc.close()
    del buf
    osc = orc.OracleScene.from_scene_data(sd)
    acc = None if accum0 is None else np.array(accum0, np.float32, copy=True)
    for k in range(frames):
        g = f.copy()
        g.frame_index = f.frame_index + k
        g.reset_accum = f.reset_accum if k == 0 else 0
        acc, _ = osc.render(g, 32, 0, accum=acc)
    return got, acc, st


def _check(label, got, ref):
    same = float(np.mean(np.all(got == ref, axis=-1)))
    print(f"{label}: equal to the oracle on {same:.6f} of pixels", flush=True)
    assert np.array_equal(got, ref), label


def _with_lights(lights, seed=23):
    sd = dmt.generate(0, seed=seed, width=48, height=36)
    sd.lights = np.array(lights, np.float32).reshape(-1, 20)
    return sd


def _unit(v):
    v = np.asarray(v, np.float64)
    return tuple(v / np.linalg.norm(v))


POINT = scenes.point_light((1.0, 6.0, 2.0), (30.0, 27.0, 22.0))
SPOT = scenes.light(1, (0.0, 7.0, 0.5), (60.0, 60.0, 50.0), dir=_unit((0.2, -1.0, 0.1)), cos_outer=0.80,
                    cos_inner=0.92)
SPHERE = scenes.sphere_light((-1.5, 6.5, 1.0), 0.6, (6.0, 7.0, 8.0))
QUAD = scenes.quad_light((0.0, 8.0, 0.0), (0.0, -1.0, 0.0), (1.5, 0.0, 0.0), 1.5, (10.0, 10.0, 10.0))
TRI = scenes.triangle_light((-3.0, 6.0, -1.0), (-2.0, 6.0, 0.5), (-3.5, 6.2, 0.5), (9.0, 6.0, 4.0))

LIGHT_SETS = [("point", [POINT]), ("spot", [SPOT]), ("sphere", [SPHERE]),
              ("mixed tree", [POINT, SPOT, SPHERE, QUAD, TRI])]


@pytest.mark.parametrize("idx", range(len(LIGHT_SETS)), ids=[s[0] for s in LIGHT_SETS])
def test_light_types_equal_the_oracle(device, idx):
    label, lights = LIGHT_SETS[idx]
    sd = _with_lights(lights)
    f = sd.frame.copy()
    f.spp = 2
    got, ref, st = _both(device, sd, f)
    _check(label, got, ref)
    assert st["shadow_rays"] > 0 and got[..., :3].max() > 0.0
    if len(lights) > 1:   # uniform pick instead of the tree
        g = f.copy()
        g.light_tree_enabled = 0
        got, ref, _ = _both(device, sd, g)
        _check(label + ", light tree off", got, ref)


def test_ema_accumulation_equals_the_oracle(device):
    sd = _with_lights([POINT, QUAD])
    f = sd.frame.copy()
    f.ema_alpha = 0.5
    got, ref, _ = _both(device, sd, f, frames=4)
    _check("EMA 0.5, 4 frames", got, ref)
    assert np.all(got[..., 3] == 2.0)   # the count is capped at 1 / (1 - alpha)
    # the batched kernel accumulates the same way
    sc = dmt.Scene(device, sd)
    b = dmt.Batch(device, [sc], [f])
    b.render(0, 4)
    device.synchronize()
    assert np.array_equal(sc.readback(f.width, f.height), ref)
    b.close()
    sc.close()


def test_sanitising_of_frame_values_and_history(device):
    """An emitter of 3e38 seen by every sample of a 2-spp pixel sums to inf
    (frame value 0); an emitter with a negative channel gives negative frame
    values (clamped to 0); NaN and inf in the previous accumulation restart
    those pixels."""
    huge = scenes.sphere((-1.0, 0.0, 0.0), 0.8, albedo=(0, 0, 0), emission=(3e38, 1.0, 1.0))
    neg = scenes.sphere((1.2, 0.0, 0.0), 0.8, albedo=(0, 0, 0), emission=(-2.0, 0.5, 1.0))
    floor = scenes.plane((0, 1, 0), 1.0, albedo=(0.5, 0.5, 0.5))
    sd = scenes.scene([huge, neg, floor], [POINT], env=np.full((8, 16, 3), 0.2, np.float32))
    f = scenes.frame((0.0, 0.5, 5.0), (0.0, 0.0, 0.0), 40.0, 40, 24, spp=2, bounces=3)
    acc0 = np.zeros((24, 40, 4), np.float32)
    acc0[..., :3] = 0.25
    acc0[..., 3] = 3.0
    acc0[2:6, 3:9, 0] = np.nan
    acc0[10:12, 20:30, 3] = np.inf
    acc0[15, 5:35, 1] = -np.inf
    f.reset_accum = 0
    got, ref, _ = _both(device, sd, f, frames=2, accum0=acc0)
    _check("sanitising", got, ref)
    assert np.all(np.isfinite(got))
    assert np.all(got[2:6, 3:9, 3] == 2.0) and np.all(got[10:12, 20:30, 3] == 2.0)   # restarted
    assert got[0, 0, 3] == 5.0
    assert np.all(got[..., :3] >= 0.0)
    # the huge emitter's pixels: every frame value there was inf -> 0
    g = f.copy()
    g.reset_accum = 1
    one, ref1, _ = _both(device, sd, g, frames=1)
    _check("sanitising, one frame", one, ref1)
    c = (12, 12)   # the huge sphere fills this pixel: red sums to inf, green and blue to 1
    assert one[c][0] == 0.0 and one[c][1] == 1.0 and one[c][2] == 1.0


def test_full_batched_workload_equals_per_scene_renders(device):
    """The 64-scene batched workload (16.2 M pixels, 18.8 M triangles): two
    frames of the batch (graph replay) equal two frames of each scene alone."""
    scs, frames, sds = [], [], []
    for i in range(64):
        sd = dmt.generate_batch_scene(1, i)
        scs.append(dmt.Scene(device, sd))
        frames.append(sd.frame.copy())
    print("64 scenes built", flush=True)
    b = dmt.Batch(device, scs, frames)
    b.render(0, 2)
    device.synchronize()
    imgs = [sc.readback(f.width, f.height) for sc, f in zip(scs, frames)]
    b.close()
    for i, (sc, f) in enumerate(zip(scs, frames)):
        sc.render_frames(f, 2)
        want = sc.readback(f.width, f.height)
        assert np.array_equal(imgs[i], want), i
    for sc in scs:
        sc.close()


def test_full_differentiable_workload_against_finite_differences(device):
    sd = dmt.generate(4, seed=1)
    f = sd.frame.copy()
    assert (f.width, f.height, f.spp, f.max_bounces, sd.prims.shape[0]) == (512, 512, 16, 4, 257)
    f.flags = dmt.DMT_FRAME_NO_RUSSIAN_ROULETTE
    f.reset_accum = 1
    sc = dmt.Scene(device, sd)
    gt = sc.materials()
    sc.render(f)
    target = sc.readback(512, 512)[..., :3].astype(np.float64)
    rng = np.random.default_rng(101)
    init = gt.copy()
    init[1:] = gt[1:] + rng.normal(0.0, 0.1, size=gt[1:].shape).astype(np.float32)
    init[1:, 0:3] = np.clip(init[1:, 0:3], 0.05, 1.0)
    init[1:, 4:7] = np.maximum(init[1:, 4:7], 0.02)
    init = init.astype(np.float32)

    def loss(m):
        sc.set_materials(m)
        sc.render(f)
        img = sc.readback(512, 512)[..., :3].astype(np.float64)
        return float(((img - target) ** 2).sum()), img

    _, img = loss(init)
    dimg = torch.zeros((512, 512, 4), device="cuda", dtype=torch.float32)
    dimg[..., :3] = torch.from_numpy((2.0 * (img - target)).astype(np.float32)).cuda()
    grads = torch.zeros((257, dmt.MATERIAL_FLOATS), device="cuda", dtype=torch.float32)
    torch.cuda.synchronize()
    sc.set_materials(init)
    sc.adjoint(f, dimg.data_ptr(), grads.data_ptr())
    device.synchronize()
    g = grads.cpu().numpy().astype(np.float64)
    kinds = [0, 1, 2, 4, 5, 6]
    big = np.abs(g[1:, kinds]).max()
    cand = [(p, k) for p in range(1, 257) for k in kinds if abs(g[p, k]) >= 0.01 * big]
    pick = [cand[i] for i in rng.choice(len(cand), size=min(24, len(cand)), replace=False)]
    h = 1e-3
    bad = []
    for p, k in pick:
        mp, mm = init.copy(), init.copy()
        mp[p, k] += h
        mm[p, k] -= h
        fd = (loss(mp)[0] - loss(mm)[0]) / (2.0 * h)
        rel = abs(g[p, k] - fd) / abs(fd)
        print(f"prim {p} param {k}: adjoint {g[p, k]:+.6e} fd {fd:+.6e} rel {rel:.2e}", flush=True)
        if rel > 1e-3:
            bad.append((p, k, rel))
    assert not bad, bad
    sc.close()


def _convergence(device, sd, spp, gpu_frames, ref_frames):
    f = sd.frame.copy()
    f.spp = spp
    W, H = f.width, f.height
    sc = dmt.Scene(device, sd)
    vals = []
    for k in range(gpu_frames):
        g = f.copy()
        g.frame_index, g.reset_accum = k, 1
        sc.render(g)
        vals.append(sc.readback(W, H)[..., :3].astype(np.float64))
    sc.close()
    V = np.array(vals)
    mean, var = V.mean(0), V.var(0, ddof=1)
    osc = orc.OracleScene.from_scene_data(sd)
    ref = np.zeros_like(mean)
    for k in range(ref_frames):
        g = f.copy()
        g.frame_index = 1_000_000 + k   # independent of the GPU frames' streams
        ref += osc.render_f64(g)
    ref /= ref_frames
    expect = float((var * (1.0 / gpu_frames + 1.0 / ref_frames)).sum())
    got = float(((mean - ref) ** 2).sum())
    sigma = math.sqrt(float(var.mean()) * (1.0 / gpu_frames + 1.0 / ref_frames) / var.size)
    dm = float(mean.mean() - ref.mean())
    print(f"{W}x{H} x {spp} spp: squared difference {got:.4e} vs noise {expect:.4e} (ratio {got / expect:.3f}); "
          f"image means differ by {dm:+.3e} = {dm / sigma:+.2f} sigma", flush=True)
    assert got <= 1.3 * expect
    assert abs(dm) <= 5.0 * sigma


def test_gpu_converges_to_the_fp64_oracle_spheres(device):
    _convergence(device, dmt.generate(0, seed=1, width=64, height=48), 32, 256, 16)


def test_gpu_converges_to_the_fp64_oracle_many_lights(device):
    sd = dmt.generate(2, seed=3, spheres=256, subdivisions=4, random_tris=5000, lights=256, width=64, height=36)
    _convergence(device, sd, 16, 256, 16)
