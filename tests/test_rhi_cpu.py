"""The RHI adapter without a GPU: PtPush decoding (field offsets, feature
gates, short pushes), the host-side resource rules the software device sets
(WriteBuffer range checks, BLAS validation, TLAS with unknown BLAS), the
exported C++ interface, and the reference-side binding compiled against the
reference's own RHI headers."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

import _dmt_rhi as rhi
import dmt
from rhi_engine import NO_TEX_TILE, OFF, PTPUSH_SIZE, pack_ptpush

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"


def _frame():
    sd = dmt.generate(0, seed=2, width=64, height=48)
    f = sd.frame.copy()
    f.frame_index = 17
    f.reset_accum = 1
    f.spp = 3
    f.max_bounces = 5
    f.refract_bounces = 6
    f.firefly_clamp = 7.5
    f.ema_alpha = 0.25
    f.mis_enabled = 0
    f.light_tree_enabled = 1
    return f


def test_ptpush_size_matches_the_engine_layout():
    # the engine asserts these offsets beside its PtPush (Engine.cpp:9409-9423)
    assert PTPUSH_SIZE == 1792
    assert (OFF["sun_and_mode"], OFF["accum_params"], OFF["hosek_params"], OFF["tonemap_params"]) == (240, 784, 1184, 1600)


def test_decode_reads_every_frame_field():
    f = _frame()
    push = pack_ptpush(f, prim_count=65, light_count=3, env_present=True, env_intensity=2.5, tlas_present=True,
                       tonemap_op=2)
    d = rhi.decode_pathtrace_push(push)
    assert d["pos_fovtan"] == pytest.approx(list(f.pos_fovtan), abs=0)
    assert d["fwd_aspect"] == pytest.approx(list(f.fwd_aspect), abs=0)
    assert d["right_xyz"] == pytest.approx(list(f.right_xyz), abs=0)
    assert d["up_xyz"] == pytest.approx(list(f.up_xyz), abs=0)
    assert (d["frame_index"], d["spp"], d["max_bounces"], d["reset_accum"]) == (17, 3, 5, 1)
    assert (d["mis_enabled"], d["light_tree_enabled"], d["refract_bounces"]) == (0, 1, 6)
    assert (d["firefly_clamp"], d["ema_alpha"], d["flags"]) == (7.5, 0.25, 0)
    assert (d["prim_count"], d["light_count"], d["env_map_present"], d["tlas_present"]) == (65, 3, 1, 1)
    assert (d["env_intensity"], d["tonemap_op"], d["motion_blur"], d["ignored"]) == (2.5, 2, 0, "")


def test_decode_refract_budget_and_light_tree_gate():
    f = _frame()
    for budget, want in ((0.0, 0), (0.5, 0), (1.0, 1), (4.9, 4), (16.0, 16)):
        push = pack_ptpush(f, prim_count=1, light_count=1, env_present=True,
                           overrides={"w2j_row1": ("4f", (0.0, 1.0, 0.0, budget))})
        assert rhi.decode_pathtrace_push(push)["refract_bounces"] == want
    # the engine clears light_tree_enabled when the cvar is off or there are no nodes (Engine.cpp:8133-8137)
    push = pack_ptpush(f, prim_count=1, light_count=0, env_present=True)
    assert rhi.decode_pathtrace_push(push)["light_tree_enabled"] == 0


GATES = [
    ("sun_and_mode", "4f", (0.0, -0.5, -0.86, 2.0), "procedural sky"),
    ("sun_and_mode", "4f", (0.0, -0.5, -0.86, 3.0), "Hosek-Wilkie"),
    ("hdri_lights_count", "I", 9, "more than 8 HDRI cluster lights"),
    ("vol_params", "4f", (0.01, 0.0, 1.0, 8.0), "participating media"),
    ("clouds_p1", "4f", (100.0, 200.0, 0.5, 0.2), "participating media"),
    ("clouds_p3", "4f", (0.0, 0.0, 0.0, 1.0), "participating media"),
    ("sdf_params", "4I", (2, 128, 1, 0), "SDF iteration heat map"),
    ("restir_enabled", "I", 1, "ReSTIR"),
    ("mat_override_count", "4I", (1, 0, 0, 0), "material overrides"),
    ("smoke_params", "4f", (1.0, 0.0, 0.0, 0.0), "smoke volumes"),
    ("write_shadow_vis", "I", 1, "SIGMA shadow visibility"),
    ("ocean_params0", "4f", (1.0, 256.0, 0.0, 0.0), "FFT ocean"),
]


@pytest.mark.parametrize("field,fmt,value,what", GATES, ids=[g[0] + "-" + g[3] for g in GATES])
def test_strict_refuses_and_lenient_lists_each_feature(field, fmt, value, what):
    f = _frame()
    push = pack_ptpush(f, prim_count=1, light_count=1, env_present=True, tlas_present=True,
                       overrides={field: (fmt, value)})
    with pytest.raises(ValueError, match=what):
        rhi.decode_pathtrace_push(push, strict=True)
    d = rhi.decode_pathtrace_push(push, strict=False)
    assert what in d["ignored"]
    assert d["spp"] == 3   # the frame itself still decodes


def test_decode_reads_the_camera_sample_features():
    """dof_params 320, dof_bokeh_params 336, spectral_params 352, motion_blur_params
    1024, mesh_motion_prev/curr 1088/1104 (Engine.cpp:8725-8739, 9020-9023, 9061-9068)
    land in the frame; the gates are the shader's (> 0 aperture, > 0.5 flags)."""
    f = _frame()
    push = pack_ptpush(f, prim_count=1, light_count=1, env_present=True, overrides={
        "dof_params": ("4f", (0.08, 6.5, 6.0, 0.25)), "dof_bokeh_params": ("4f", (0.4, 0.0, 0.0, 0.0)),
        "spectral_params": ("4f", (1.0, 0.0042, 0.0, 0.0)), "motion_blur_params": ("4f", (1.0, 0.016, 0.0, 0.0)),
        "mesh_motion_prev": ("4f", (0.5, -1.0, 2.0, 1.0)), "mesh_motion_curr": ("4f", (0.75, -1.0, 2.5, 0.0))})
    d = rhi.decode_pathtrace_push(push, strict=True)
    assert d["ignored"] == ""
    assert np.allclose(d["dof"], [0.08, 6.5, 6.0, 0.25]) and np.isclose(d["cat_eye"], 0.4)
    assert (d["spectral"], d["motion_blur"], d["mesh_shift_active"]) == (1, 1, 1)
    assert np.isclose(d["cauchy_c"], 0.0042)
    assert np.allclose(d["mesh_shift_prev"], [0.5, -1.0, 2.0]) and np.allclose(d["mesh_shift_curr"], [0.75, -1.0, 2.5])
    push = pack_ptpush(f, prim_count=1, light_count=1, env_present=True, overrides={
        "spectral_params": ("4f", (0.5, 0.0042, 0.0, 0.0)), "motion_blur_params": ("4f", (0.5, 0.016, 0.0, 0.0)),
        "mesh_motion_prev": ("4f", (0.5, -1.0, 2.0, 0.5))})
    d = rhi.decode_pathtrace_push(push)
    assert (d["spectral"], d["motion_blur"], d["mesh_shift_active"]) == (0, 0, 0)


def test_decode_reads_the_sdf_fields():
    """sdf_params 816 (count, r_sdf_max_iters) and sdf_params_f 832 (r_sdf_epsilon)
    (Engine.cpp:7947-7981) reach the frame; the fractal knobs (.w, f.yzw) are
    read by fractal clusters only, which the scene build refuses."""
    f = _frame()
    push = pack_ptpush(f, prim_count=1, light_count=1, env_present=True, overrides={
        "sdf_params": ("4I", (3, 96, 0, 12)), "sdf_params_f": ("4f", (2.5e-4, 8.0, 1.0, 0.0))})
    d = rhi.decode_pathtrace_push(push, strict=True)
    assert d["ignored"] == ""
    assert (d["sdf_count"], d["sdf_max_iters"]) == (3, 96) and np.isclose(d["sdf_epsilon"], 2.5e-4)


def test_decode_reads_the_denoiser_fields():
    f = _frame()
    vp = tuple(float(i) * 0.5 for i in range(16))
    pv = tuple(float(i) * -0.25 for i in range(16))
    push = pack_ptpush(f, prim_count=1, light_count=1, env_present=True, overrides={
        "denoiser_enabled": ("I", 1), "write_normal_gbuffer": ("I", 1), "write_albedo_gbuffer": ("I", 0),
        "write_specular_albedo_gbuffer": ("I", 1), "write_roughness_gbuffer": ("I", 1),
        "write_specular_hit_distance_gbuffer": ("I", 0), "write_hdr_aux": ("I", 1),
        "w2j_row0": ("4f", (1.0, 0.0, 0.0, 0.0)), "curr_view_proj": ("16f", vp), "prev_view_proj": ("16f", pv)})
    d = rhi.decode_pathtrace_push(push)
    assert (d["denoiser"], d["write_normal"], d["write_albedo"], d["write_specular_albedo"]) == (1, 1, 0, 1)
    assert (d["write_roughness"], d["write_specular_distance"], d["write_hdr_aux"], d["hdr_pipeline"]) == (1, 0, 1, 0)
    assert d["curr_view_proj"] == list(vp) and d["prev_view_proj"] == list(pv)
    # the per-frame colour input is the mean only when the frame resets it
    g = f.copy()
    g.reset_accum = 0
    push = pack_ptpush(g, prim_count=1, light_count=1, env_present=True,
                       overrides={"denoiser_enabled": ("I", 1), "write_hdr_aux": ("I", 1)})
    with pytest.raises(ValueError, match="without reset_accum"):
        rhi.decode_pathtrace_push(push)
    assert "without reset_accum" in rhi.decode_pathtrace_push(push, strict=False)["ignored"]
    # without the denoiser the colour input is the running mean: no reset needed
    push = pack_ptpush(g, prim_count=1, light_count=1, env_present=True, overrides={"write_hdr_aux": ("I", 1)})
    assert rhi.decode_pathtrace_push(push)["write_hdr_aux"] == 1


def test_decode_reads_the_sky_lights_and_the_denoiser_jitter():
    """sun_and_mode.xyz 240, sun_extra 480, sun_extra2.x 1616, moon_dir_phase
    448, moon_extra 464 (the gradient sky's sun and moon, PathTrace.slang:3112,
    3212, 6546-6667); hdri_lights_dir 496, hdri_lights_col 624, count 752 (the
    env map's cluster lights, :6503); halton_jitter 96 while the denoiser runs
    (:5730).  None of them is refused."""
    f = _frame()
    sun, moon = (0.3, 0.8, -0.52, 0.0), (-0.4, 0.6, 0.69, 2.1)
    push = pack_ptpush(f, prim_count=1, light_count=1, env_present=False,
                       overrides={"sun_and_mode": ("4f", sun[:3] + (1.0,)), "sun_extra": ("4f", (1.5, 1.01, 0.9, 6.4e6)),
                                  "sun_extra2": ("4f", (0.7, 0.0, 0.0, 0.0)), "moon_dir_phase": ("4f", moon),
                                  "moon_extra": ("4f", (2.0, 0.97, 0.8, 1.2))})
    d = rhi.decode_pathtrace_push(push)
    assert d["ignored"] == "" and d["env_map_present"] == 0
    assert np.allclose(d["sun"], sun) and np.allclose(d["moon"], moon)
    assert np.allclose(d["sun_extra"], (1.5, 1.01, 0.9, 0.0)) and np.allclose(d["moon_extra"], (2.0, 0.97, 0.8, 0.0))
    assert abs(d["sun_disc_brightness"] - 0.7) < 1e-7
    # cluster lights count with the env map only (PathTrace.slang:3680)
    dirs = [(0.0, 1.0, 0.0, 0.25), (0.6, 0.8, 0.0, 0.75)]
    cols = [(10.0, 9.0, 8.0, 0.0), (1.0, 2.0, 3.0, 0.0)]
    ov = {"hdri_lights_count": ("I", 2), "env_total_luminance": ("f", 12.5)}
    for i in range(2):
        ov[f"hdri_dir{i}"] = ("4f", dirs[i])
        ov[f"hdri_col{i}"] = ("4f", cols[i])
    d = rhi.decode_pathtrace_push(pack_ptpush(f, prim_count=1, light_count=1, env_present=True, overrides=ov))
    assert d["hdri_light_count"] == 2 and d["ignored"] == "" and d["env_total_luminance"] == 12.5
    assert np.allclose(np.reshape(d["hdri_dir"], (8, 4))[:2], dirs)
    assert np.allclose(np.reshape(d["hdri_col"], (8, 4))[:2], cols)
    assert not np.any(np.reshape(d["hdri_dir"], (8, 4))[2:])
    d = rhi.decode_pathtrace_push(pack_ptpush(f, prim_count=1, light_count=1, env_present=False, overrides=ov))
    assert d["hdri_light_count"] == 0
    # the denoiser's Halton offset replaces sample 0's jitter
    ov = {"denoiser_enabled": ("I", 1), "halton_jitter": ("2f", (0.25, -0.125))}
    d = rhi.decode_pathtrace_push(pack_ptpush(f, prim_count=1, light_count=1, env_present=True, overrides=ov))
    assert d["flags"] & 2 and d["jitter"] == [0.25, -0.125]
    d = rhi.decode_pathtrace_push(pack_ptpush(f, prim_count=1, light_count=1, env_present=True,
                                              overrides={"halton_jitter": ("2f", (0.25, -0.125))}))
    assert not d["flags"] & 2


def test_short_push_is_refused_and_missing_fields_read_as_zero():
    f = _frame()
    full = pack_ptpush(f, prim_count=9, light_count=1, env_present=True)
    for n in (0, 64, 111):
        with pytest.raises(ValueError, match="shorter than the 112 B"):
            rhi.decode_pathtrace_push(full[:n])
    d = rhi.decode_pathtrace_push(full[:112])   # the head only (SoftwareTracer.cpp:38-56)
    assert (d["prim_count"], d["spp"], d["env_map_present"]) == (9, 3, 1)
    assert (d["light_count"], d["mis_enabled"], d["light_tree_enabled"], d["refract_bounces"]) == (0, 0, 0, 0)
    assert (d["firefly_clamp"], d["ema_alpha"], d["tonemap_op"]) == (0.0, 0.0, 0)
    # a head-only push without the env map has no sun or moon (zero directions)
    noenv = pack_ptpush(f, prim_count=9, light_count=1, env_present=False)
    d = rhi.decode_pathtrace_push(noenv[:112])
    assert d["sun"] == [0.0] * 4 and d["moon"] == [0.0] * 4 and d["ignored"] == ""
    # bytes past the 2048-byte push budget are dropped, not read
    assert rhi.decode_pathtrace_push(full + bytes(4096))["prim_count"] == 9


def test_host_resources_without_a_device():
    d = rhi.HipDevice(strict=True)
    if not d.is_device_lost():
        pytest.skip("a HIP device is present: covered by the GPU tests")
    assert "dmt_device_create" in d.last_error()
    assert d.supports_hardware_rt()
    # buffers are host copies: written and read back, out-of-range writes refused
    b = d.create_buffer(64)
    assert b
    d.write_buffer(b, np.arange(8, dtype=np.float32), 16)
    raw = d.readback_buffer(b, 64)
    assert np.array_equal(np.frombuffer(raw, np.float32)[4:12], np.arange(8, dtype=np.float32))
    d.write_buffer(b, np.zeros(8, np.float32), 40)
    assert "exceeds buffer size 64" in d.last_error()
    assert np.array_equal(np.frombuffer(d.readback_buffer(b, 64), np.float32), np.frombuffer(raw, np.float32))
    assert d.readback_buffer(b, 65) is None
    # BLAS: empty meshes and out-of-range indices are refused
    pos = np.zeros((3, 3), np.float32)
    assert d.create_blas(pos, np.array([0, 1, 2], np.uint32))
    assert d.create_blas(np.zeros((0, 3), np.float32), np.zeros(0, np.uint32)) == 0
    assert d.create_blas(pos, np.array([0, 1, 3], np.uint32)) == 0 and "out of range" in d.last_error()
    assert d.create_blas(pos, np.array([0, 1], np.uint32)) == 0
    # TLAS: unknown BLAS ids are skipped, an empty instance list gives no TLAS
    blas = d.create_blas(pos, np.array([0, 1, 2], np.uint32))
    assert d.create_tlas([(blas, None, 0, 0xFF), (999999, None, 0, 0xFF)])
    assert d.create_tlas([]) == 0
    with pytest.raises(ValueError):
        d.create_tlas([(blas, [1.0] * 11, 0, 0xFF)])
    # textures need device memory
    assert d.create_texture(4, 4) == 0
    assert d.allocated_bytes() > 0
    d.close()


def _nm_demangled(path):
    if not shutil.which("nm"):
        pytest.skip("nm not available")
    out = subprocess.run(["nm", "-DC", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return out


def test_adapter_exports_the_device_interface():
    lib = os.path.join(ROOT, "demont-engine_amd", "libdmt_rhi.so")
    syms = _nm_demangled(lib)
    methods = ["CreateBuffer", "CreateTexture", "CreateComputePipeline", "CreateBLAS", "CreateTLAS", "DestroyBuffer",
               "DestroyTexture", "DestroyPipeline", "DestroyAccelStruct", "WriteBuffer", "WriteTexture", "BeginFrame",
               "EndFrame", "AcquireCommandBuffer", "Submit", "WaitIdle", "Resize", "ReadbackTexture", "ReadbackBuffer",
               "DeviceName", "CurrentAllocatedBytes", "IsDeviceLost", "LastError", "ActiveScene", "MeshRefits"]
    for m in methods:
        assert f"dmt::rhi::HipDevice::{m}(" in syms or f"dmt::rhi::HipDevice::{m}[abi:" in syms, m
    for m in ["BindComputePipeline", "BindBuffer", "BindStorageTexture", "BindAccelStruct", "PushConstants",
              "Dispatch", "CopyBufferToTexture", "ClearStorageTexture", "Barrier"]:
        assert f"dmt::rhi::HipCommandBuffer::{m}(" in syms, m
    assert "dmt::rhi::DecodePathtracePush(" in syms
    assert "dmt::rhi::CreateHipDevice(" in syms


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF_SRC, "rhi")), reason="reference RHI headers not present")
def test_reference_side_binding_compiles_against_the_reference_headers(tmp_path):
    """pt::rhi::HipDevice (csrc/rhi/pt_rhi_hip_device.cpp) implements every pure
    virtual of the reference's Device and CommandBuffer and links against
    libdmt_rhi.so with no undefined symbol."""
    src = os.path.join(ROOT, "demont-engine_amd", "csrc", "rhi", "pt_rhi_hip_device.cpp")
    obj, so = str(tmp_path / "binding.o"), str(tmp_path / "binding.so")
    subprocess.run(["g++", "-std=c++2b", "-Wall", "-Wextra", "-Werror", "-fPIC", "-c", src, "-I", REF_SRC, "-I",
                    os.path.join(ROOT, "include"), "-o", obj], check=True)
    libdir = os.path.join(ROOT, "demont-engine_amd")
    subprocess.run(["g++", "-shared", "-o", so, obj, "-L", libdir, "-ldmt_rhi", "-Wl,--no-undefined",
                    "-Wl,-rpath," + libdir], check=True)
    assert "pt::rhi::CreateHipDevice(" in _nm_demangled(so)


def test_decode_reads_the_mesh_map_tiles():
    """mesh_tex_indices (1152, Engine.cpp:9131-9134) into frame.mesh_tex, with
    DMT_FRAME_MESH_TEXTURES when any channel names a tile."""
    f = _frame()
    push = pack_ptpush(f, prim_count=1, light_count=1, env_present=True, tlas_present=True,
                       overrides={"mesh_tex_indices": ("4I", (3, NO_TEX_TILE, 5, 0))})
    d = rhi.decode_pathtrace_push(push, strict=True)
    assert d["mesh_tex"] == [3, NO_TEX_TILE, 5, 0]
    assert d["flags"] & dmt.FRAME_MESH_TEXTURES
    assert d["ignored"] == ""
    d = rhi.decode_pathtrace_push(pack_ptpush(f, prim_count=1, light_count=1, env_present=True, tlas_present=True),
                                  strict=True)
    assert d["mesh_tex"] == [NO_TEX_TILE] * 4 and not (d["flags"] & dmt.FRAME_MESH_TEXTURES)


def test_decode_reads_the_water_fields():
    """water_params0 912 (absorption rgb, index), water_params1 928 (wave scale,
    amplitude, speed, time), water_state 1040 (camera in water, deep tint),
    as Engine.cpp:8883-8890 and :9263-9266 fill them."""
    f = _frame()
    ov = {"water_params0": ("4f", (0.1, 0.2, 0.4, 1.33)), "water_params1": ("4f", (1.5, 0.3, 0.7, 12.5)),
          "water_state": ("4f", (1.0, 0.01, 0.02, 0.05))}
    d = rhi.decode_pathtrace_push(pack_ptpush(f, prim_count=1, light_count=1, env_present=True, overrides=ov),
                                  strict=True)
    assert d["water_absorb"] == pytest.approx([0.1, 0.2, 0.4]) and d["water_ior"] == pytest.approx(1.33)
    assert d["wave"] == pytest.approx([1.5, 0.3, 0.7, 12.5])
    assert d["camera_in_water"] == 1 and d["deep_color"] == pytest.approx([0.01, 0.02, 0.05])
    ov["water_state"] = ("4f", (0.4, 0.0, 0.0, 0.0))   # the engine writes 0.0 or 1.0; > 0.5 is in the water
    d = rhi.decode_pathtrace_push(pack_ptpush(f, prim_count=1, light_count=1, env_present=True, overrides=ov),
                                  strict=True)
    assert d["camera_in_water"] == 0
