"""Pins the oracle's scalar building blocks (CPU only).

* sin/cos/acos/atan2: the fp32 Cephes restatement against numpy's float64
  functions, within 3 ulp over the argument ranges the path tracer uses.
* pcgHash / randf (shaders/PathTraceMath.slang:16-24): the PCG hash of
  Jarzynski & Olano, "Hash Functions for GPU Rendering" (JCGT 2020) -- an LCG
  step followed by the RXS-M-XS output permutation -- evaluated here in
  uint64 numpy and compared with the oracle through its rendered RNG use.
"""
import numpy as np
import pytest

import oracle_ctypes as orc


def _ulp_err(got, ref):
    got = got.astype(np.float64)
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    ulp = np.maximum(ulp, np.spacing(np.float32(1e-30)))
    return np.abs(got - ref) / ulp


@pytest.mark.parametrize("fn,lo,hi", [("sin", -7.0, 7.0), ("cos", -7.0, 7.0)])
def test_cephes_sincos(fn, lo, hi):
    x = np.linspace(lo, hi, 200001, dtype=np.float32)
    got = orc.math_f32(fn, x)
    ref = getattr(np, fn)(x.astype(np.float64))
    # absolute error bound near zeros of the function, ulp bound elsewhere
    err = np.abs(got.astype(np.float64) - ref)
    big = np.abs(ref) > 1e-3
    assert np.max(_ulp_err(got[big], ref[big])) <= 3.0
    assert np.max(err[~big]) <= 2e-7


def test_cephes_acos():
    x = np.linspace(-1.0, 1.0, 200001, dtype=np.float32)
    got = orc.math_f32("acos", x)
    ref = np.arccos(x.astype(np.float64))
    big = ref > 1e-3
    assert np.max(_ulp_err(got[big], ref[big])) <= 3.0
    assert np.max(np.abs(got[~big] - ref[~big])) <= 2e-7
    assert got[0] == np.float32(np.pi) and got[-1] == 0.0


def test_cephes_atan2_quadrants():
    rng = np.random.default_rng(7)
    y = rng.normal(size=200000).astype(np.float32)
    x = rng.normal(size=200000).astype(np.float32)
    got = orc.math_f32("atan2", y, x)
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    big = np.abs(ref) > 1e-3
    # the y/x quotient is rounded before the atan polynomial: one more ulp
    assert np.max(_ulp_err(got[big], ref[big])) <= 4.0
    # axis special cases
    ys = np.array([0, 1, -1, 0, 0], np.float32)
    xs = np.array([1, 0, 0, -1, 0], np.float32)
    assert np.allclose(orc.math_f32("atan2", ys, xs), [0, np.pi / 2, -np.pi / 2, np.pi, 0], atol=1e-7)


def pcg_hash_np(x):
    """JCGT 2020 pcg_hash in uint64 arithmetic masked to 32 bits."""
    x = np.asarray(x, np.uint64)
    m = np.uint64(0xFFFFFFFF)
    state = (x * np.uint64(747796405) + np.uint64(2891336453)) & m
    word = (((state >> ((state >> np.uint64(28)) + np.uint64(4))) ^ state) * np.uint64(277803737)) & m
    return ((word >> np.uint64(22)) ^ word) & m


def test_pcg_properties():
    xs = np.arange(0, 1 << 16, dtype=np.uint64)
    h = pcg_hash_np(xs)
    # a bijection on the sampled range: no collisions among 65536 inputs
    assert np.unique(h).size == xs.size
    # randf = (h & 0xFFFFFF) / 2^24 is uniform in [0, 1)
    u = (h & np.uint64(0xFFFFFF)).astype(np.float64) / 16777216.0
    assert u.min() >= 0.0 and u.max() < 1.0
    counts, _ = np.histogram(u, bins=64, range=(0, 1))
    expected = xs.size / 64
    chi2 = float(((counts - expected) ** 2 / expected).sum())
    assert chi2 < 120.0   # 63 dof, p ~ 1e-5


def test_pcg_drives_the_oracle_camera_jitter():
    """The first two draws of pixel (x, y) at frame f jitter its camera ray
    (PathTrace.slang:5457, 5733).  A 1x1-pixel render with a 1-bounce miss
    into a gradient sky lets the sampled ray direction be recovered from the
    radiance; it must match the jitter predicted by pcg_hash_np."""
    import dmt
    import scenes
    sd = scenes.scene(env=None)
    for fidx in (0, 1, 17):
        f = scenes.frame((0, 0, 0), (0, 0, -1), 90.0, 1, 1, bounces=1, frame_index=fidx)
        img, _ = orc.OracleScene.from_scene_data(sd).render(f, 64, 1)
        seed = pcg_hash_np(np.uint64(fidx * 26699))
        s1 = pcg_hash_np(seed)
        s2 = pcg_hash_np(s1)
        jx = float(s1 & np.uint64(0xFFFFFF)) / 16777216.0 - 0.5
        jy = float(s2 & np.uint64(0xFFFFFF)) / 16777216.0 - 0.5
        uy = -(((0.5 + jy) / 1.0) * 2.0 - 1.0)
        ux = ((0.5 + jx) / 1.0) * 2.0 - 1.0
        d = np.array([ux, uy, -1.0])   # tan(45 deg) = 1, aspect 1
        d /= np.linalg.norm(d)
        t = min(max(0.5 * (d[1] + 1.0), 0.0), 1.0)
        sky = np.array([0.5, 0.55, 0.62]) + (np.ones(3) - np.array([0.5, 0.55, 0.62])) * t
        assert np.allclose(img[0, 0, :3], sky, rtol=1e-6), (fidx, img[0, 0, :3], sky)
