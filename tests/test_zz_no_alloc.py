"""dmt_trace_closest_noalloc, the allocation-free closest-hit query, against
dmt_trace_closest (the allocating one) through the C ABI: two calls in a row
on caller-owned device memory, the second with fewer rays, bit for bit equal
and writing nothing past its n records; a scene whose trees are deeper than
the static scratch allows is refused with DMT_ERR_LIMIT before any work."""
import numpy as np
import pytest
import torch

import dmt

pytestmark = pytest.mark.gpu

DMT_ERR_LIMIT = -7


def _scene():
    """Spheres, the ground plane and a quad light over a small mesh."""
    m = dmt.generate(1, seed=42, subdivisions=4, random_tris=3000, width=40, height=30)
    s = dmt.generate(0, seed=19, width=8, height=8)
    return dmt.SceneData(prims=s.prims, lights=s.lights, positions=m.positions, indices=m.indices, env=m.env,
                         env_intensity=m.env_intensity, frame=m.frame.copy())


def _rays(rng, n):
    o = rng.uniform(-7.0, 7.0, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # a quarter aimed at the mesh centre, so most of those hit it
    k = n // 4
    t = -o[:k] + rng.uniform(-1.0, 1.0, (k, 3)).astype(np.float32)
    d[:k] = t / np.linalg.norm(t, axis=1, keepdims=True)
    return o, d


def test_noalloc_query_equals_the_allocating_query(device):
    sc = dmt.Scene(device, _scene())
    rng = np.random.default_rng(7)
    cap = 6000
    out = torch.empty((cap, 8), dtype=torch.int32, device="cuda")
    for n in (6000, 2500):   # the second call has fewer rays
        o, d = _rays(rng, n)
        want = sc.trace(o, d)
        rays = np.zeros((n, 8), np.float32)
        rays[:, 0:3], rays[:, 4:7] = o, d
        d_rays = torch.from_numpy(rays).cuda()
        out.fill_(-1)   # sentinel: 0xFFFFFFFF in every word
        torch.cuda.synchronize()   # the rays are complete before the call
        print(f"noalloc query: {n} rays", flush=True)
        st = sc.trace_device_noalloc(d_rays.data_ptr(), n, out.data_ptr())
        assert st == 0, dmt.last_error()
        device.synchronize()
        got = out.cpu().numpy()
        assert got[:n].tobytes() == want.tobytes()
        assert np.all(got[n:] == -1)   # nothing written past the n records
        kinds = want["hit"]
        assert (kinds == 1).sum() > 100 and (kinds == 2).sum() > 100 and (kinds == 0).sum() > 100
    sc.close()


def test_noalloc_query_refuses_trees_deeper_than_its_scratch(device):
    # a chain of triangles at geometrically growing positions: the SAH build
    # peels them off one by one, 22 levels deep
    pos, idx = [], []
    for k in range(90):
        x = 1.5 ** k
        s = 0.25 * x
        b = len(pos) // 3
        pos += [x, 0, 0, x + s, 0, 0, x, s, 0]
        idx += [b, b + 1, b + 2]
    sd = dmt.SceneData(prims=np.zeros((0, 28), np.float32), lights=np.zeros((0, 20), np.float32),
                       positions=np.array(pos, np.float32).reshape(-1, 3),
                       indices=np.array(idx, np.uint32).reshape(-1, 3), env=None, env_intensity=1.0)
    sc = dmt.Scene(device, sd)
    assert sc.stats()["bvh_depth"] > (dmt.DMT_NOALLOC_SPILL_DEPTH + 12) // 3
    d_rays = torch.zeros((4, 8), dtype=torch.float32, device="cuda")
    d_out = torch.zeros((4, 8), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    st = sc.trace_device_noalloc(d_rays.data_ptr(), 4, d_out.data_ptr())
    assert st == DMT_ERR_LIMIT
    assert "depth" in dmt.last_error()
    sc.close()
