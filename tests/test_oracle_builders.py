"""Pins the host builders (CPU only): env CDF, light tree, BVHs.

Follows the reference's own unit tests for these builders:
  tests/hdrimage_cdf_test.cpp        (4x4 hand-computed CDF :233-291, single
                                      bright pixel :302-348, zero image :356-370,
                                      dark rows :381-416, sin(theta) weighting
                                      :425-466, chi^2 sampling :487-560, closed-form
                                      pdf :580-629)
  tests/pt_renderer_light_tree_test.cpp (2N-1 nodes :116-117, reachability
                                      :195-228, coincident lights :230-246,
                                      AABB/intensity invariants :253-276, PMF
                                      sums to 1 :352-391, brighter favoured :393-422)
  tests/pt_renderer_triangle_bvh_test.cpp (malformed indices :324, parent
                                      contains children :563+)
and, where oracle/_ref is built, compares node for node with the reference's
own compiled builders.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import dmt
import oracle_ctypes as orc
import scenes

REF_SO = os.path.join(os.path.dirname(orc.LIB_PATH), "_ref", "libref.so")


def ref_lib():
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    L = C.CDLL(REF_SO)
    u32p = C.POINTER(C.c_uint32)
    fp = C.POINTER(C.c_float)
    L.ref_tri_bvh.argtypes = [fp, C.c_uint32, u32p, C.c_uint32, u32p, u32p]
    L.ref_tri_bvh.restype = C.c_uint32
    L.ref_analytic_bvh.argtypes = [fp, C.c_uint32, u32p, u32p]
    L.ref_analytic_bvh.restype = C.c_uint32
    L.ref_light_tree.argtypes = [fp, C.c_uint32, fp]
    L.ref_light_tree.restype = C.c_uint32
    return L


# ---------------------------------------------------------------- env CDF

def _greyscale(W, H, f):
    rgb = np.zeros((H, W, 3), np.float32)
    for v in range(H):
        for u in range(W):
            rgb[v, u, :] = f(u, v)
    return rgb


@pytest.mark.parametrize("builder", ["oracle", "product"])
def test_env_cdf_4x4_hand_computed(builder):
    W = H = 4
    rgb = _greyscale(W, H, lambda u, v: (v + 1) / H)
    sin_t = [math.sin(math.pi * (v + 0.5) / H) for v in range(H)]
    row = [W * ((v + 1) / H) * sin_t[v] for v in range(H)]
    total_ref = sum(row)
    marg_ref = np.cumsum(row) / total_ref
    marg_ref[-1] = 1.0
    cond_ref = np.tile((np.arange(W) + 1) / W, (H, 1))
    m, c, total = orc.env_cdf(rgb) if builder == "oracle" else dmt.build_env_cdf(rgb)
    assert total == pytest.approx(total_ref, rel=1e-6)
    assert np.allclose(m, marg_ref, rtol=1e-5)
    assert np.allclose(c, cond_ref, rtol=1e-5)
    assert m[-1] == 1.0


def test_env_cdf_edge_cases_and_product_bit_exact():
    rng = np.random.default_rng(3)
    cases = {
        "lognormal": np.exp(rng.normal(size=(32, 64, 3))).astype(np.float32),
        "zero": np.zeros((8, 16, 3), np.float32),
        "single": np.zeros((8, 16, 3), np.float32),
        "dark_rows": np.exp(rng.normal(size=(16, 16, 3))).astype(np.float32),
    }
    cases["single"][5, 9, :] = 7.0
    cases["dark_rows"][3:9] = 0.0
    for name, img in cases.items():
        m, c, t = orc.env_cdf(img)
        m2, c2, t2 = dmt.build_env_cdf(img)
        assert np.array_equal(m, m2) and np.array_equal(c, c2) and t == t2, name
        assert np.all(np.diff(m) >= 0), name
        assert np.all(np.diff(c, axis=1) >= 0), name
        if name == "zero":
            assert t == 0.0 and not m.any() and not c.any()
        else:
            assert m[-1] == 1.0
        if name == "single":
            assert np.all(m[:5] == 0.0) and np.all(m[5:] == 1.0)
            assert np.all(c[5, :9] == 0.0) and np.all(c[5, 9:] == 1.0)
        if name == "dark_rows":
            assert np.all(c[3:9] == 0.0)
            assert np.all(np.diff(m)[2:8] == 0.0)


def test_env_cdf_sin_weighting():
    """Uniform image: the marginal is the normalised prefix sum of sin(theta)."""
    H, W = 32, 8
    m, c, t = orc.env_cdf(np.ones((H, W, 3), np.float32))
    s = np.sin(np.pi * (np.arange(H) + 0.5) / H)
    assert np.allclose(m, np.cumsum(s) / s.sum(), atol=2e-7)
    assert t == pytest.approx(W * s.sum(), rel=1e-12)


def _env_scene(img):
    return orc.OracleScene(np.zeros((0, 28), np.float32), np.zeros((0, 20), np.float32),
                           np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32), env=img)


def test_env_sampling_chi2_and_closed_form_pdf():
    rng = np.random.default_rng(11)
    H, W = 8, 16
    img = np.exp(rng.normal(size=(H, W, 3))).astype(np.float32)
    sc = _env_scene(img)
    n = 200000
    d, col, pdf = sc.env_sample(rng.random((n, 2), dtype=np.float32))
    th = np.arccos(np.clip(d[:, 1], -1, 1))
    ph = np.arctan2(d[:, 2], d[:, 0])
    vi = np.clip((th / np.pi * H).astype(int), 0, H - 1)
    ui = np.clip(((ph + np.pi) / (2 * np.pi) * W).astype(int), 0, W - 1)
    lum = 0.2126 * img[..., 0] + 0.7152 * img[..., 1] + 0.0722 * img[..., 2]
    sin_t = np.sin(np.pi * (np.arange(H) + 0.5) / H)[:, None]
    p = (lum * sin_t).astype(np.float64)
    p /= p.sum()
    counts = np.zeros((H, W))
    np.add.at(counts, (vi, ui), 1)
    exp = p * n
    chi2 = float(((counts - exp) ** 2 / exp).sum())
    assert chi2 < 200.0   # 127 dof
    # closed-form solid-angle pdf (hdrimage_cdf_test.cpp:580-629)
    total = float((lum.astype(np.float64) * sin_t).sum())
    pdf_ref = lum[vi, ui] * W * H / (2 * np.pi ** 2 * total)
    assert np.allclose(pdf, pdf_ref, rtol=2e-5)
    assert np.allclose(col, img[vi, ui], rtol=0)
    # pdfEnvMapDir of the sampled direction returns the same density
    assert np.allclose(sc.env_pdf(d), pdf, rtol=2e-5)
    # the density integrates to 1 over the sphere (Monte Carlo, uniform directions)
    u = rng.normal(size=(400000, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    est = float(np.mean(sc.env_pdf(u.astype(np.float32)))) * 4 * np.pi
    assert est == pytest.approx(1.0, abs=0.02)


def test_env_zero_total_disables_sampling():
    sc = _env_scene(np.zeros((4, 8, 3), np.float32))
    _, _, pdf = sc.env_sample(np.array([[0.3, 0.7]], np.float32))
    assert pdf[0] == 0.0
    assert sc.env_pdf(np.array([[0, 0.5, 0.866]], np.float32))[0] == 0.0


# ------------------------------------------------------------ light tree

def random_lights(n, seed, kinds=(0, 1, 2, 3)):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        k = kinds[i % len(kinds)]
        pos = rng.uniform(-10, 10, 3)
        I = rng.uniform(0.1, 5.0, 3)
        if k == 0:
            out.append(scenes.point_light(pos, I))
        elif k == 1:
            d = rng.normal(size=3); d /= np.linalg.norm(d)
            out.append(scenes.light(1, pos, I, dir=d, cos_outer=0.7, cos_inner=0.9))
        elif k == 2:
            out.append(scenes.sphere_light(pos, rng.uniform(0.05, 1.0), I))
        elif k == 3:
            d = rng.normal(size=3); d /= np.linalg.norm(d)
            u = np.cross(d, [0.3, 0.5, 0.81]); u /= np.linalg.norm(u)
            out.append(scenes.quad_light(pos, d, u * rng.uniform(0.1, 1), rng.uniform(0.1, 1), I))
        else:
            out.append(scenes.triangle_light(pos, pos + rng.uniform(-1, 1, 3), pos + rng.uniform(-1, 1, 3), I))
    return np.array(out, np.float32)


def _walk(nodes):
    leaves = []
    st = [0]
    while st:
        i = st.pop()
        nd = nodes[i]
        cnt = int(nd[7:8].view(np.uint32)[0])
        lf = int(nd[3:4].view(np.uint32)[0])
        if cnt:
            leaves.append(lf)
        else:
            for ch in (lf, lf + 1):
                c = nodes[ch]
                assert np.all(nd[0:3] <= c[0:3]) and np.all(nd[4:7] >= c[4:7])
                st.append(ch)
            l, r = nodes[lf], nodes[lf + 1]
            assert nd[12] == np.float32(l[12] + r[12])
    return leaves


@pytest.mark.parametrize("n", [1, 2, 3, 7, 64, 200, 500, 1000])
def test_light_tree_structure(n):
    L = random_lights(n, seed=n, kinds=(0, 1, 2, 3, 4))
    nodes = orc.light_tree(L)
    assert len(nodes) == 2 * n - 1
    leaves = _walk(nodes)
    assert sorted(leaves) == list(range(n))
    # the product builder is node-for-node identical
    assert np.array_equal(nodes.view(np.uint32), dmt.build_light_tree(L).view(np.uint32))


def test_triangle_light_leaf_summary():
    """Emissive triangle leaf (light type 4, no reference counterpart): the
    triangle's bounds, its normal as a hemisphere cone, luminance x area."""
    a, b, c = np.array([0.0, 3.0, 0.0]), np.array([1.0, 3.0, 0.0]), np.array([0.0, 3.0, 2.0])
    Le = np.array([2.0, 1.0, 0.5], np.float32)
    nodes = orc.light_tree(np.array([scenes.triangle_light(a, b, c, Le)], np.float32))
    nd = nodes[0]
    assert np.allclose(nd[0:3], [0, 3, 0]) and np.allclose(nd[4:7], [1, 3, 2])
    n = np.cross(b - a, c - a); n /= np.linalg.norm(n)
    assert np.allclose(nd[8:11], n) and nd[11] == 0.0
    lum = 0.2126 * 2.0 + 0.7152 * 1.0 + 0.0722 * 0.5
    assert abs(nd[12] - lum * 1.0) < 1e-5   # area = |(1,0,0) x (0,0,2)| / 2 = 1
    assert np.array_equal(nodes.view(np.uint32), dmt.build_light_tree(
        np.array([scenes.triangle_light(a, b, c, Le)], np.float32)).view(np.uint32))


def test_light_tree_coincident_lights():
    L = np.array([scenes.point_light((1, 2, 3), (1, 1, 1))] * 9, np.float32)
    nodes = orc.light_tree(L)
    assert len(nodes) == 17
    assert sorted(_walk(nodes)) == list(range(9))


def test_light_tree_matches_reference_builder():
    R = ref_lib()
    for n, seed in ((1, 1), (5, 2), (64, 3), (1000, 4)):
        L = random_lights(n, seed, kinds=(0, 1, 2, 3))
        out = np.zeros((2 * n - 1, 16), np.float32)
        k = R.ref_light_tree(L.ctypes.data_as(C.POINTER(C.c_float)), n, out.ctypes.data_as(C.POINTER(C.c_float)))
        assert k == 2 * n - 1
        mine = orc.light_tree(L)
        assert np.array_equal(mine.view(np.uint32), out.view(np.uint32)), f"n={n}"


def test_light_tree_pmf_sums_to_one_and_favours_bright():
    L = random_lights(64, seed=9, kinds=(0, 2))
    sd = scenes.scene(lights=L, env=scenes.black_env())
    sc = orc.OracleScene.from_scene_data(sd)
    rng = np.random.default_rng(5)
    for _ in range(20):
        p = rng.uniform(-12, 12, 3).astype(np.float32)
        n = rng.normal(size=3).astype(np.float32)
        n /= np.linalg.norm(n)
        pmf = sc.light_pmf(p, n, 64)
        assert float(pmf.astype(np.float64).sum()) == pytest.approx(1.0, abs=1e-5)
        assert np.all(pmf >= 0)
    # two clusters, one 100x brighter: the bright one receives most of the mass
    lights = [scenes.point_light((5 + 0.1 * i, 0, 0), (100, 100, 100)) for i in range(4)] + \
             [scenes.point_light((-5 - 0.1 * i, 0, 0), (1, 1, 1)) for i in range(4)]
    sc2 = orc.OracleScene.from_scene_data(scenes.scene(lights=lights, env=scenes.black_env()))
    pmf = sc2.light_pmf(np.array([0, 3, 0], np.float32), np.array([0, -1, 0], np.float32), 8)
    assert pmf[:4].sum() > 0.95


# ------------------------------------------------------------------ BVHs

def _tri_soup(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-5, 5, (n, 1, 3))
    v = (c + rng.uniform(-0.3, 0.3, (n, 3, 3))).reshape(-1, 3).astype(np.float32)
    return v, np.arange(3 * n, dtype=np.uint32).reshape(n, 3)


def test_median_tri_bvh_matches_reference_builder():
    R = ref_lib()
    for n, seed in ((1, 1), (5, 2), (1000, 3), (20000, 4)):
        v, idx = _tri_soup(n, seed)
        nodes, perm = orc.tri_bvh(v, idx)
        rn = np.zeros((2 * n + 1, 8), np.uint32)
        rp = np.zeros(n, np.uint32)
        k = R.ref_tri_bvh(v.ctypes.data_as(C.POINTER(C.c_float)), v.shape[0],
                          idx.ctypes.data_as(C.POINTER(C.c_uint32)), idx.size,
                          rn.ctypes.data_as(C.POINTER(C.c_uint32)), rp.ctypes.data_as(C.POINTER(C.c_uint32)))
        assert k == len(nodes)
        assert np.array_equal(nodes, rn[:k]) and np.array_equal(perm[:n], rp), f"n={n}"


def test_median_tri_bvh_skips_malformed_triangles():
    v, idx = _tri_soup(10, 5)
    idx = idx.copy()
    idx[3, 1] = 10_000   # out of range: dropped, as TriangleBvh.cpp:49 does
    nodes, perm = orc.tri_bvh(v, idx)
    assert 3 not in set(perm[:9].tolist())


def test_product_bvh4_invariants():
    v, idx = _tri_soup(5000, 8)
    nodes, order, depth = dmt.build_bvh4(v, idx)
    assert sorted(order.tolist()) == list(range(5000))
    f = nodes.view(np.float32).reshape(-1, 32)
    u = nodes.view(np.uint32).reshape(-1, 32)
    tri_lo = v[idx].min(axis=1)
    tri_hi = v[idx].max(axis=1)
    seen = np.zeros(5000, int)
    max_d = 0
    st = [(0, 1)]
    while st:
        ni, d = st.pop()
        max_d = max(max_d, d)
        lo = np.stack([f[ni, 0:4], f[ni, 8:12], f[ni, 16:20]], 1)
        hi = np.stack([f[ni, 4:8], f[ni, 12:16], f[ni, 20:24]], 1)
        for k in range(4):
            ch = int(u[ni, 24 + k])
            if ch == 0xFFFFFFFF:
                continue
            if ch & 0x80000000:
                first, cnt = ch & ((1 << 27) - 1), ((ch >> 27) & 15) + 1
                assert cnt <= 4
                tris = order[first:first + cnt]
                seen[tris] += 1
                assert np.all(tri_lo[tris] >= lo[k]) and np.all(tri_hi[tris] <= hi[k])
            else:
                clo = np.stack([f[ch, 0:4], f[ch, 8:12], f[ch, 16:20]], 1)
                chi = np.stack([f[ch, 4:8], f[ch, 12:16], f[ch, 20:24]], 1)
                valid = u[ch, 24:28] != 0xFFFFFFFF
                assert np.all(clo[valid] >= lo[k]) and np.all(chi[valid] <= hi[k])
                st.append((ch, d + 1))
    assert np.all(seen == 1)
    assert max_d == depth


def test_oracle_bvh_traversal_equals_bruteforce():
    """orc_trace (median BVH, robust slab test) == orc_trace_bruteforce, bit
    for bit, on the scene kinds of the configs -- the closest-hit equivalence
    the reference checks with 1000 rays (pt_renderer_triangle_bvh_test.cpp:510-556)."""
    rng = np.random.default_rng(21)
    for sd in (dmt.generate(0, seed=2), dmt.generate(1, seed=2, subdivisions=3, random_tris=3000),
               dmt.generate(4, seed=3, spheres=200)):
        sc = orc.OracleScene.from_scene_data(sd)
        o = rng.uniform(-6, 6, (3000, 3)).astype(np.float32)
        d = rng.normal(size=(3000, 3))
        d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
        a, b = sc.trace(o, d), sc.trace(o, d, brute=True)
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
        assert (a["hit"] > 0).mean() > 0.1
