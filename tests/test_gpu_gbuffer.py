"""Denoiser inputs on the GPU: dmt_render_gbuffer equals the oracle's G-buffer
pass bit for bit (every output, at a pixel stride that leaves the other floats
untouched), and through the RHI device an engine frame with the denoiser on
fills the bound G-buffer textures and the denoiser's colour input."""
import numpy as np
import pytest
import torch

import _dmt_rhi as rhi
import dmt
import oracle_ctypes as orc
from rhi_engine import RGBA32F, STORAGE, EngineSim, tonemap_reference, view_proj

pytestmark = pytest.mark.gpu

# output -> (first channel in the oracle's 16-float record, floats compared, floats written)
OUTS = {"depth": (0, 1, 1), "motion": (1, 2, 2), "normal": (3, 3, 4), "albedo": (6, 3, 4),
        "specular_albedo": (9, 3, 4), "roughness": (12, 1, 1), "specular_hit_distance": (13, 1, 1)}


def _scenes():
    out = [("spheres+plane", dmt.generate(0, seed=41, width=56, height=40))]
    out.append(("mesh + env", dmt.generate(1, seed=42, subdivisions=4, random_tris=2000, width=64, height=36)))
    out.append(("many lights", dmt.generate(2, seed=43, spheres=150, subdivisions=3, random_tris=3000, lights=64,
                                            width=48, height=32)))
    return out


SCENES = _scenes()


def _gpu_gbuffer(device, sd, frame, vc, vp, stride):
    n = frame.width * frame.height
    bufs = {k: torch.full((n * stride,), -7.0, dtype=torch.float32, device="cuda:0")
            for k in list(OUTS) + ["cloud_transmittance"]}
    torch.cuda.synchronize()
    sc = dmt.Scene(device, sd)
    sc.gbuffer(frame, vc, vp, stride, **{k: v.data_ptr() for k, v in bufs.items()})
    device.synchronize()
    out = {k: v.cpu().numpy().reshape(frame.height, frame.width, stride) for k, v in bufs.items()}
    sc.close()
    return out


@pytest.mark.parametrize("idx", range(len(SCENES)), ids=[s[0] for s in SCENES])
@pytest.mark.parametrize("stride", [4, 6])
def test_gbuffer_equals_the_oracle(device, idx, stride):
    label, sd = SCENES[idx]
    f = sd.frame
    vc = view_proj(f).astype(np.float32)
    shift = np.array(f.right_xyz[:3]) * 0.2 - np.array(f.fwd_aspect[:3]) * 0.3
    vp = view_proj(f, pos=np.array(f.pos_fovtan[:3]) + shift).astype(np.float32)
    got = _gpu_gbuffer(device, sd, f, vc, vp, stride)
    want = orc.OracleScene.from_scene_data(sd).gbuffer(f, vc, vp)
    for k, (c0, nc, nw) in OUTS.items():
        g = got[k]
        same = float(np.mean(np.all(g[..., :nc] == want[..., c0:c0 + nc], axis=-1)))
        print(f"{label} stride {stride} {k}: equal to the oracle on {same:.6f} of pixels")
        assert np.array_equal(g[..., :nc], want[..., c0:c0 + nc]), k
        if nw == 4:
            assert np.all(g[..., 3] == 0.0)
        assert np.all(g[..., nw:] == -7.0), k          # the rest of each pixel's stride untouched
    assert np.all(got["cloud_transmittance"][..., 0] == 1.0)
    assert (want[..., 0] < 1e9).any()


def test_gbuffer_argument_checks(device):
    label, sd = SCENES[0]
    f = sd.frame
    sc = dmt.Scene(device, sd)
    buf = torch.zeros(f.width * f.height * 4, dtype=torch.float32, device="cuda:0")
    M = view_proj(f).astype(np.float32)
    with pytest.raises(dmt.DmtError):
        sc.gbuffer(f, M, M, 3, normal=buf.data_ptr())      # normal writes 4 floats
    with pytest.raises(dmt.DmtError):
        sc.gbuffer(f, M, M, 4, depth=buf.data_ptr() + 2)   # not float-aligned
    sc.gbuffer(f, M, M, 1)                                  # nothing to write: a no-op
    sc.close()


def test_engine_denoiser_frame_fills_the_bound_textures(device):
    label, sd = SCENES[1]
    f = sd.frame.copy()
    f.reset_accum = 1
    w, h = f.width, f.height
    dev = rhi.HipDevice()
    eng = EngineSim(dev, sd)
    tex = {slot: dev.create_texture(w, h, RGBA32F, STORAGE) for slot in (2, 3, 4, 8, 9, 10, 11, 12, 13)}
    vc = view_proj(f).astype(np.float32)
    vp = view_proj(f, pos=np.array(f.pos_fovtan[:3]) + np.array(f.up_xyz[:3]) * 0.25).astype(np.float32)
    flags = {"denoiser_enabled": ("I", 1), "write_normal_gbuffer": ("I", 1), "write_albedo_gbuffer": ("I", 1),
             "write_specular_albedo_gbuffer": ("I", 1), "write_roughness_gbuffer": ("I", 1),
             "write_specular_hit_distance_gbuffer": ("I", 1), "write_hdr_aux": ("I", 1),
             "curr_view_proj": ("16f", tuple(float(v) for v in vc)),
             "prev_view_proj": ("16f", tuple(float(v) for v in vp))}
    for k in range(2):   # the engine resets the accumulation every frame while a denoiser runs
        g = f.copy()
        g.frame_index = f.frame_index + k
        eng.frame(g, textures=tex, overrides=dict(flags, w2j_row0=("4f", (1.0, 0.0, 0.0, 1.0))))
        assert dev.last_error() == "", dev.last_error()
    acc = eng.accum_image()
    assert np.all(acc[..., 3] == 1.0)
    want = orc.OracleScene.from_scene_data(sd).gbuffer(f, vc, vp)
    read = {slot: dev.readback_texture(t) for slot, t in tex.items()}
    for slot, (c0, nc) in {3: (0, 1), 4: (1, 2), 8: (3, 3), 9: (6, 3), 11: (9, 3), 12: (12, 1), 13: (13, 1)}.items():
        assert np.array_equal(read[slot][..., :nc], want[..., c0:c0 + nc]), slot
    assert np.all(read[10][..., 0] == 1.0)
    # r_hdr_pipeline on: the denoiser colour input is the frame's radiance
    assert np.array_equal(read[2][..., :3], acc[..., :3]) and np.all(read[2][..., 3] == 1.0)
    # r_hdr_pipeline off: tone-mapped, no transfer function
    g = f.copy()
    g.frame_index = f.frame_index + 2
    eng.frame(g, textures=tex, tonemap_op=1, overrides=dict(flags, w2j_row0=("4f", (1.0, 0.0, 0.0, 0.0))))
    acc = eng.accum_image()
    col = dev.readback_texture(tex[2])
    err = float(np.max(np.abs(col[..., :3] - tonemap_reference(acc[..., :3], 1.0, 1, oetf=False))))
    print(f"tone-mapped denoiser colour input: max error {err:.2e}")
    assert err <= 2e-5
    # a G-buffer texture of another size is refused before any work
    small = dev.create_texture(w - 1, h, RGBA32F, STORAGE)
    before = eng.accum_image()
    eng.frame(g, textures={**tex, 3: small}, overrides=dict(flags, w2j_row0=("4f", (1.0, 0.0, 0.0, 1.0))))
    assert "slot 3" in dev.last_error()
    assert np.array_equal(eng.accum_image(), before)
    dev.close()
