"""The PtPush byte offsets the RHI device decodes, checked against the
reference engine's own declaration (CPU only; needs the reference tree, so it
is skipped where that is absent).

The engine declares `struct PtPush` locally in its frame function
(src/engine/Engine.cpp:7243-7820) with plain float / std::uint32_t members
and two array bounds from src/engine/Engine.h (kMaxHdriLights,
kMaxMatOverrides).  The test reads that declaration as text, compiles it in a
temporary directory next to include/dmt_rhi.h, and static-asserts that
offsetof(PtPush, field) equals every ptpush::k* constant DecodePathtracePush
reads, that sizeof(PtPush) is ptpush::kSize and that the 112-byte head ends
where the view-projection matrices start.  The compiled program then prints
every field's offset, and the engine-call replay's encoder table
(tests/rhi_engine.py OFF) must agree with it field for field.
"""
import os
import re
import subprocess

import pytest

REF = "/root/reference/src/engine"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# decoder constant -> reference field
DECODER_FIELDS = {
    "kPosFovTan": "pos_fovtan", "kFwdAspect": "fwd_aspect", "kRightXyz": "right_xyz", "kUpXyz": "up_xyz",
    "kFrameIndex": "frame_index", "kResetAccum": "reset_accum", "kMaxBounces": "max_bounces",
    "kTlasPresent": "tlas_present", "kPrimCount": "prim_count", "kSpp": "spp",
    "kDenoiserEnabled": "denoiser_enabled", "kEnvMapPresent": "env_map_present",
    "kEnvIntensity": "env_intensity", "kCurrViewProj": "curr_view_proj", "kPrevViewProj": "prev_view_proj",
    "kSunAndMode": "sun_and_mode", "kW2jRow0": "w2j_row0", "kW2jRow1": "w2j_row1", "kW2jRow2": "w2j_row2",
    "kDofParams": "dof_params", "kDofBokehParams": "dof_bokeh_params", "kSpectralParams": "spectral_params",
    "kVolParams": "vol_params", "kCloudsP1": "clouds_p1", "kCloudsP3": "clouds_p3",
    "kMoonDirPhase": "moon_dir_phase", "kHdriLightsCount": "hdri_lights_count",
    "kWriteNormalGbuffer": "write_normal_gbuffer", "kWriteAlbedoGbuffer": "write_albedo_gbuffer",
    "kWriteHdrAux": "write_hdr_aux", "kMisEnabled": "mis_enabled", "kAccumParams": "accum_params",
    "kSdfParams": "sdf_params", "kSdfParamsF": "sdf_params_f", "kWriteShadowVis": "write_shadow_vis", "kLightCount": "light_count",
    "kLightTreeEnabled": "light_tree_enabled", "kWriteSpecularAlbedo": "write_specular_albedo_gbuffer",
    "kWriteRoughness": "write_roughness_gbuffer", "kWriteSpecularHitDistance": "write_specular_hit_distance_gbuffer",
    "kRestirEnabled": "restir_enabled", "kSmokeParams": "smoke_params", "kMotionBlurParams": "motion_blur_params",
    "kCloudsRuntime": "clouds_runtime", "kMeshMotionPrev": "mesh_motion_prev", "kMeshMotionCurr": "mesh_motion_curr",
    "kMeshTexIndices": "mesh_tex_indices", "kMatOverrideCount": "mat_override_count",
    "kTonemapParams": "tonemap_params", "kWaterParams0": "water_params0", "kWaterParams1": "water_params1",
    "kWaterState": "water_state", "kOceanParams0": "ocean_params0",
}


def _reference_struct():
    with open(os.path.join(REF, "Engine.cpp")) as fh:
        text = fh.read()
    start = text.index("    struct PtPush {")
    end = text.index("    } push{};", start)
    body = text[start:end] + "    };\n"
    with open(os.path.join(REF, "Engine.h")) as fh:
        hdr = fh.read()
    consts = {}
    for name in ("kMaxHdriLights", "kMaxMatOverrides"):
        m = re.search(r"static\s+constexpr\s+std::uint32_t\s+" + name + r"\s*=\s*(\d+)\s*;", hdr)
        assert m, name
        consts[name] = int(m.group(1))
    fields = re.findall(r"^\s*(?:float|std::uint32_t|std::int32_t)\s+(\w+)\s*[\[;]", body, re.M)
    return body, consts, fields


@pytest.fixture(scope="module")
def layout(tmp_path_factory):
    if not os.path.isfile(os.path.join(REF, "Engine.cpp")):
        pytest.skip("reference engine sources not present")
    body, consts, fields = _reference_struct()
    missing = sorted(set(DECODER_FIELDS.values()) - set(fields))
    assert not missing, f"decoder fields absent from the reference PtPush: {missing}"
    asserts = "\n".join(
        f'static_assert(offsetof(PtPush, {f}) == dmt::rhi::ptpush::{k}, "{k} != offsetof({f})");'
        for k, f in DECODER_FIELDS.items())
    prints = "\n".join(f'    std::printf("{f} %zu\\n", offsetof(PtPush, {f}));' for f in fields)
    src = f"""#include <cstddef>
#include <cstdint>
#include <cstdio>
#include "dmt_rhi.h"
struct Engine {{
    static constexpr std::uint32_t kMaxHdriLights = {consts['kMaxHdriLights']};
    static constexpr std::uint32_t kMaxMatOverrides = {consts['kMaxMatOverrides']};
}};
{body}
{asserts}
static_assert(sizeof(PtPush) == dmt::rhi::ptpush::kSize, "kSize != sizeof(PtPush)");
static_assert(offsetof(PtPush, curr_view_proj) == dmt::rhi::ptpush::kHeadBytes, "head size");
int main() {{
    std::printf("sizeof %zu\\n", sizeof(PtPush));
{prints}
    return 0;
}}
"""
    d = tmp_path_factory.mktemp("ptpush")
    cpp, exe = d / "ptpush.cpp", d / "ptpush"
    cpp.write_text(src)
    r = subprocess.run(["g++", "-std=c++2b", "-Wall", "-I", os.path.join(ROOT, "include"), str(cpp), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    offs = {}
    for line in out.splitlines():
        k, v = line.split()
        offs[k] = int(v)
    return offs


def test_decoder_offsets_match_the_reference_declaration(layout):
    # the static_asserts above already compiled; the size is the 2048-byte
    # software-device push limit's tenant
    assert layout["sizeof"] <= 2048
    assert layout["sizeof"] == 1792


def test_engine_replay_encoder_matches_the_reference_declaration(layout):
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    rhi_engine = pytest.importorskip("rhi_engine")
    bad = {k: (v, layout.get(k)) for k, v in rhi_engine.OFF.items() if layout.get(k) != v}
    assert not bad, f"tests/rhi_engine.py OFF differs from the reference layout: {bad}"
    assert rhi_engine.PTPUSH_SIZE == layout["sizeof"]
