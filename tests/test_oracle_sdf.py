"""Pins of the oracle's SDF clusters (CPU only): the shape distances and
normals, the CSG ops, the sphere trace, the cluster order in traceScene, and
a rendered closed form.  The expected values come from geometry computed
here in float64, independently of the distance formulas:
  * distances: the box by clamping (outside |p - clamp(p, -h, h)|, inside
    minus the nearest face gap), the rounded box as the box's outer parallel
    surface, the torus and the capsule by dense sampling of their core circle
    and segment, the plane by its equation;
  * normals: central differences of those distances;
  * ops (SdfPrimitives.slang:276-301): k = 0 gives the set algebra (union min,
    subtraction max(a, -b), intersection max); at equal arguments the smooth
    blend is h = 1/2 and adds -k/4 (union) or +k/4; displace adds its constant;
    an op of the procedural set (compiled out of the reference's default
    build) evaluates to 1e30;
  * sphere trace (SdfPrimitives.slang:1330-1444): against a ray-sphere root
    and a box face the hit lies at most eps / cos(angle) before the surface
    and never past it; a ray from inside a sphere leaves it at r; the cluster
    box clips the trace; two steps reach a sphere 1e-3 inside its box, so a
    cap of one step misses;
  * traceScene (PathTrace.slang:2212-2306): clusters after the analytic
    primitives, strictly closer to replace a hit, the lower index at a tie;
  * a Lambert SDF slab under a point light receives albedo/pi I cos / d^2.
"""
import math

import numpy as np

import oracle_ctypes as orc
import scenes
from scenes import (SDF_BOX, SDF_CAPSULE, SDF_DISPLACE, SDF_INTERSECT, SDF_LEAF, SDF_PLANE, SDF_ROUND_BOX,
                    SDF_SPHERE, SDF_SUBTRACT, SDF_TORUS, SDF_UNION, sdf_ball, sdf_box, sdf_cluster, sdf_node)

BIG = (-10, -10, -10), (10, 10, 10)


def _leaf(shape, params, center=(0, 0, 0)):
    return sdf_cluster([sdf_node(SDF_LEAF, shape, params=params, center=center)], *BIG)


def _box_dist(p, h):
    p, h = np.asarray(p, np.float64), np.asarray(h, np.float64)
    q = np.clip(p, -h, h)
    out = np.linalg.norm(p - q, axis=-1)
    inside = np.all(np.abs(p) <= h, axis=-1)
    gap = np.min(h - np.abs(p), axis=-1)
    return np.where(inside, -gap, out)


def _core_dist(p, pts):
    """min distance from each p to a dense point set (a curve)."""
    d = np.full(p.shape[0], np.inf)
    for k in range(0, pts.shape[0], 4096):
        blk = pts[k:k + 4096]
        d = np.minimum(d, np.min(np.linalg.norm(p[:, None, :] - blk[None, :, :], axis=-1), axis=1))
    return d


def _points(n=400, seed=3, lo=-0.8, hi=0.8):
    return np.random.default_rng(seed).uniform(lo, hi, (n, 3)).astype(np.float32)


def _num_grad(f, p, h=1e-4):
    g = np.zeros_like(p, dtype=np.float64)
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        g[:, a] = (f(p + e) - f(p - e)) / (2 * h)
    return g / np.linalg.norm(g, axis=1, keepdims=True)


def test_shape_distances_match_independent_geometry():
    p = _points()
    p64 = p.astype(np.float64)
    c = np.array([0.1, -0.05, 0.07])
    # sphere
    got = orc.sdf_eval(_leaf(SDF_SPHERE, (0.4, 0, 0, 0), c), p)[:, 0]
    assert np.allclose(got, np.linalg.norm(p64 - c, axis=1) - 0.4, atol=2e-6)
    # box, inside and outside
    h = (0.3, 0.2, 0.1)
    got = orc.sdf_eval(_leaf(SDF_BOX, (*h, 0), c), p)[:, 0]
    assert np.allclose(got, _box_dist(p64 - c, h), atol=2e-6)
    assert (got < 0).sum() >= 3
    # rounded box: the parallel surface at 0.05 outside the box
    got = orc.sdf_eval(_leaf(SDF_ROUND_BOX, (*h, 0.05), c), p)[:, 0]
    out = _box_dist(p64 - c, h) > 0
    assert np.allclose(got[out], _box_dist(p64 - c, h)[out] - 0.05, atol=2e-6)
    # torus (core circle of radius 0.5 in the xz plane), capsule (core segment y in [-0.3, 0.3])
    th = np.linspace(0, 2 * np.pi, 200000, endpoint=False)
    ring = np.stack([0.5 * np.cos(th), 0 * th, 0.5 * np.sin(th)], 1) + c
    got = orc.sdf_eval(_leaf(SDF_TORUS, (0.5, 0.1, 0, 0), c), p[:200])[:, 0]
    assert np.allclose(got, _core_dist(p64[:200], ring) - 0.1, atol=2e-5)
    seg = np.stack([0 * th[:20000], np.linspace(-0.3, 0.3, 20000), 0 * th[:20000]], 1) + c
    got = orc.sdf_eval(_leaf(SDF_CAPSULE, (0.3, 0.15, 0, 0), c), p[:200])[:, 0]
    assert np.allclose(got, _core_dist(p64[:200], seg) - 0.15, atol=2e-5)
    # plane n . (p - c) + d
    n = np.array([0.0, 0.6, 0.8])
    got = orc.sdf_eval(_leaf(SDF_PLANE, (*n, 0.25), c), p)[:, 0]
    assert np.allclose(got, (p64 - c) @ n + 0.25, atol=2e-6)
    # an unknown shape is far away
    assert np.all(orc.sdf_eval(_leaf(9, (0.1, 0, 0, 0), c), p[:4])[:, 0] == np.float32(1e30))


def test_shape_normals_match_numerical_gradients():
    p = _points(300, seed=5)
    p64 = p.astype(np.float64)
    c = np.array([0.05, 0.02, -0.04])
    h = (0.3, 0.2, 0.1)
    cases = [
        (_leaf(SDF_SPHERE, (0.4, 0, 0, 0), c), lambda q: np.linalg.norm(q - c, axis=1) - 0.4),
        (_leaf(SDF_BOX, (*h, 0), c), lambda q: _box_dist(q - c, h)),
        (_leaf(SDF_ROUND_BOX, (*h, 0.05), c), lambda q: _box_dist(q - c, h) - 0.05),
        (_leaf(SDF_TORUS, (0.5, 0.1, 0, 0), c),
         lambda q: np.hypot(np.hypot(q[:, 0] - c[0], q[:, 2] - c[2]) - 0.5, q[:, 1] - c[1]) - 0.1),
        (_leaf(SDF_CAPSULE, (0.3, 0.15, 0, 0), c),
         lambda q: np.linalg.norm(q - c - np.stack([0 * q[:, 0], np.clip(q[:, 1] - c[1], -0.3, 0.3), 0 * q[:, 0]], 1),
                                  axis=1) - 0.15),
    ]
    for cl, f in cases:
        got = orc.sdf_eval(cl, p)[:, 1:].astype(np.float64)
        assert np.allclose(np.linalg.norm(got, axis=1), 1.0, atol=1e-6)
        d = f(p64)
        ok = np.abs(d) > 0.02   # away from the medial creases, where the gradient jumps
        ok &= np.abs(np.abs(p64 - c) - np.array(h)).min(axis=1) > 0.02 if cl is cases[1][0] else ok
        want = _num_grad(f, p64)
        assert np.all(np.sum(got[ok] * want[ok], axis=1) > 1 - 1e-4)


def _two_balls(op, k, c1=(-0.2, 0, 0), r1=0.3, c2=(0.25, 0, 0), r2=0.25, extra=None):
    nodes = [sdf_node(SDF_LEAF, SDF_SPHERE, params=(r1, 0, 0, 0), center=c1),
             sdf_node(SDF_LEAF, SDF_SPHERE, params=(r2, 0, 0, 0), center=c2),
             sdf_node(op, 0, 0, 1, params=(k, 0, 0, 0))]
    return sdf_cluster(nodes, *BIG)


def test_csg_ops_and_smooth_blends():
    p = _points(300, seed=7)
    p64 = p.astype(np.float64)
    da = np.linalg.norm(p64 - np.array([-0.2, 0, 0]), axis=1) - 0.3
    db = np.linalg.norm(p64 - np.array([0.25, 0, 0]), axis=1) - 0.25
    for op, want in ((SDF_UNION, np.minimum(da, db)), (SDF_SUBTRACT, np.maximum(da, -db)),
                     (SDF_INTERSECT, np.maximum(da, db))):
        got = orc.sdf_eval(_two_balls(op, 0.0), p)[:, 0]
        assert np.allclose(got, want, atol=2e-6), op
        # far from the blend band the smooth ops are the hard ones
        k = 0.05
        sel = np.abs(da - db) > k if op != SDF_SUBTRACT else np.abs(da + db) > k
        got = orc.sdf_eval(_two_balls(op, k), p)[:, 0]
        assert np.allclose(got[sel], want[sel], atol=2e-6), op
    # on the bisector of two equal balls the union blends at h = 1/2: d - k/4
    k = 0.2
    q = np.array([[0.0, 0.1, 0.2], [0.0, -0.3, 0.05]], np.float32)
    cl = _two_balls(SDF_UNION, k, c1=(-0.3, 0, 0), r1=0.2, c2=(0.3, 0, 0), r2=0.2)
    d = np.linalg.norm(q.astype(np.float64) - np.array([0.3, 0, 0]), axis=1) - 0.2
    assert np.allclose(orc.sdf_eval(cl, q)[:, 0], d - k / 4, atol=2e-6)
    # intersection of equal arguments: d + k/4; subtraction where b = -a: a + k/4
    cl = _two_balls(SDF_INTERSECT, k, c1=(-0.3, 0, 0), r1=0.2, c2=(0.3, 0, 0), r2=0.2)
    assert np.allclose(orc.sdf_eval(cl, q)[:, 0], d + k / 4, atol=2e-6)
    cl = sdf_cluster([sdf_node(SDF_LEAF, SDF_PLANE, params=(1, 0, 0, 0)),
                      sdf_node(SDF_LEAF, SDF_PLANE, params=(1, 0, 0, 0)),
                      sdf_node(SDF_SUBTRACT, 0, 0, 1, params=(k, 0, 0, 0))], *BIG)
    x = np.array([[0.3, 0, 0]], np.float32)   # a = 0.3, b = 0.3: h = clamp(0.5 - 0.3 / k) = 0; result max(a, -b)
    assert np.allclose(orc.sdf_eval(cl, x)[:, 0], 0.3, atol=1e-6)
    x0 = np.array([[0.0, 0.4, 0.0]], np.float32)   # a = b = 0: h = 1/2, lerp(0, 0) + k/4
    assert np.allclose(orc.sdf_eval(cl, x0)[:, 0], k / 4, atol=1e-6)
    # displace adds its constant; a procedural op (6, twist) is 1e30, so a union with it is the leaf
    cl = sdf_cluster([sdf_node(SDF_LEAF, SDF_SPHERE, params=(0.3, 0, 0, 0)),
                      sdf_node(SDF_DISPLACE, 0, 0, 0, params=(-0.05, 0, 0, 0))], *BIG)
    assert np.allclose(orc.sdf_eval(cl, p)[:, 0], np.linalg.norm(p64, axis=1) - 0.35, atol=2e-6)
    cl = sdf_cluster([sdf_node(SDF_LEAF, SDF_SPHERE, params=(0.3, 0, 0, 0)),
                      sdf_node(6, 0, 0, 0, params=(1.0, 0, 0, 0)),
                      sdf_node(SDF_UNION, 0, 0, 1)], *BIG)
    assert np.allclose(orc.sdf_eval(cl, p)[:, 0], np.linalg.norm(p64, axis=1) - 0.3, atol=2e-6)
    # the union's normal is the nearer ball's
    cl = _two_balls(SDF_UNION, 0.0)
    got = orc.sdf_eval(cl, p)
    near = np.where((da < db)[:, None], p64 - [-0.2, 0, 0], p64 - [0.25, 0, 0])
    near /= np.linalg.norm(near, axis=1, keepdims=True)
    assert np.allclose(got[:, 1:], near, atol=1e-5)


def test_sphere_trace_against_exact_intersections():
    rng = np.random.default_rng(11)
    eps = 1e-4
    c, r = np.array([0.2, -0.1, 0.3]), 0.5
    cl = sdf_ball(c, r)
    o = rng.uniform(-3, 3, (500, 3))
    o = c + 2.5 * o / np.linalg.norm(o, axis=1, keepdims=True)
    tgt = c + rng.uniform(-0.3, 0.3, (500, 3))
    d = (tgt - o) / np.linalg.norm(tgt - o, axis=1, keepdims=True)
    o32, d32 = o.astype(np.float32), d.astype(np.float32)
    got = orc.sdf_trace(cl, o32, d32, 128, eps)
    oc = o32.astype(np.float64) - c
    dd = d32.astype(np.float64)
    b = np.sum(oc * dd, 1)
    t_exact = -b - np.sqrt(b * b - (np.sum(oc * oc, 1) - r * r))
    hit_p = o32 + d32 * t_exact[:, None]
    cos = np.abs(np.sum((hit_p - c) / r * dd, 1))
    assert np.all(got[:, 0] > 0)
    gap = t_exact - got[:, 0]
    assert np.all(gap > -1e-5) and np.all(gap <= eps / cos + 1e-5)
    n_want = (o32 + d32 * got[:, :1] - c) / r
    assert np.allclose(got[:, 1:], n_want / np.linalg.norm(n_want, axis=1, keepdims=True), atol=1e-5)
    # a box face at normal incidence: at most eps short, the face normal exactly
    box = sdf_box((0, 0, 0), (0.4, 0.3, 0.2))
    g = orc.sdf_trace(box, [[0.05, 0.1, 3.0]], [[0.0, 0.0, -1.0]], 128, eps)[0]
    assert 0.0 <= (3.0 - 0.2) - g[0] <= eps + 1e-6 and tuple(g[1:]) == (0.0, 0.0, 1.0)
    # from the centre the trace leaves the ball at r
    g = orc.sdf_trace(cl, [c], [[0.0, 0.6, 0.8]], 128, eps)[0]
    assert abs(g[0] - r) <= eps + 1e-6
    # the cluster box clips the trace: a box the ray misses hides the ball; a
    # box that starts inside the ball starts the trace there, which then stops
    # where the ray leaves the ball (|d| steps from inside)
    ball = [sdf_node(SDF_LEAF, SDF_SPHERE, params=(r, 0, 0, 0), center=c)]
    above = sdf_cluster(ball, c + np.array([-0.6, 0.2, -0.6]), c + 0.6)
    assert orc.sdf_trace(above, [c + [-3.0, 0, 0]], [[1.0, 0, 0]], 128, eps)[0, 0] == -1.0
    inner = sdf_cluster(ball, c + np.array([0.3, -0.6, -0.6]), c + 0.6)
    t = orc.sdf_trace(inner, [c + [-3.0, 0, 0]], [[1.0, 0, 0]], 128, eps)[0, 0]
    assert 3.0 + r - eps <= t <= 3.0 + r + eps
    # a head-on ray meets the ball 1e-3 inside its box: the second step hits, a cap of one misses
    o1, d1 = [c + [0.0, 0.0, 3.0]], [[0.0, 0.0, -1.0]]
    assert orc.sdf_trace(cl, o1, d1, 1, eps)[0, 0] == -1.0
    assert orc.sdf_trace(cl, o1, d1, 2, eps)[0, 0] > 0


def test_trace_scene_orders_clusters_after_analytic_prims():
    a = sdf_ball((0, 0, 0), 0.5, albedo=(0.9, 0.1, 0.1))
    b = sdf_ball((0, 0, 0), 0.5, albedo=(0.1, 0.9, 0.1))
    sph = scenes.sphere((0, 0, 1.0), 0.2)
    o = np.array([[0.0, 0.0, 3.0]], np.float32)
    d = np.array([[0.0, 0.0, -1.0]], np.float32)
    # two identical clusters: the lower index
    h = orc.OracleScene.from_scene_data(scenes.scene(sdf=[a, b])).trace(o, d)[0]
    assert (h["hit"], h["prim"]) == (3, 0) and 2.5 - 1e-4 <= h["t"] <= 2.5
    h = orc.OracleScene.from_scene_data(scenes.scene(sdf=[b, a])).trace(o, d)[0]
    assert (h["hit"], h["prim"]) == (3, 0)
    # an analytic sphere in front wins; behind, the cluster does
    h = orc.OracleScene.from_scene_data(scenes.scene([sph], sdf=[a])).trace(o, d)[0]
    assert h["hit"] == 1 and abs(h["t"] - 1.8) < 1e-5
    h = orc.OracleScene.from_scene_data(scenes.scene([scenes.sphere((0, 0, -1.0), 0.2)], sdf=[a])).trace(o, d)[0]
    assert (h["hit"], h["prim"]) == (3, 0)
    # the material is the cluster's
    assert h["mat"] == 0 and h["user_id"] == 0


def test_lambert_sdf_slab_under_a_point_light():
    """A Lambert SDF box slab (top face y = 0) lit by a point light: one bounce
    is deterministic, a/pi I cos / d^2 at the hit point a trace step above the face."""
    a = np.array([0.6, 0.5, 0.4])
    slab = sdf_box((0, -0.5, 0), (5.0, 0.5, 5.0), albedo=tuple(a))
    I = np.array([30.0, 30.0, 30.0])
    lp = np.array([1.0, 4.0, -0.5])
    sd = scenes.scene(lights=[scenes.point_light(lp, I)], env=scenes.black_env(), sdf=[slab])
    f = scenes.frame((0.0, 1.0, 1.5), (0.0, 0.0, 0.0), 0.02, 4, 4, bounces=1)
    acc, _ = orc.OracleScene.from_scene_data(sd).render(f, 64, 0)
    got = acc[..., :3].reshape(-1, 3).mean(0)
    v = lp - np.zeros(3)
    want = a / math.pi * I * v[1] / np.linalg.norm(v) ** 3
    assert np.allclose(got, want, rtol=2e-4), (got, want)
